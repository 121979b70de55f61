// parse.h -- the reference's run parameters (src/parse.h:10-35) plus the
// device fields added here.  defaults() sets the reference's values
// (src/parse.cxx:7-29).
#ifndef QX_SHIM_PARSE_H
#define QX_SHIM_PARSE_H

#include <string>
#include "enum.h"

struct Params {
    int N;
    int dim;
    int nthr;
    double ai;
    double Lbox;
    double dtr;
    double dt;
    int nsteps;
    int nmeas;
    double norm;
    double sol_bkg;
    bool verb;
    bool readj;
    std::string dir;
    ICType ictype;
    CosmoType cosmotype;
    FFTPlanType plan;
    MeasureType measinfo;
    // additions: device arithmetic and ordinal, k = 0 convention
    Precision precision;
    int device;
    int k0_mode;        // 1: reference (-1/S), 0: zeroed
    int rho_bar_mode;   // 0: fixed rho_mean, 1: actual mean
};

void defaults(Params &p);

#endif
