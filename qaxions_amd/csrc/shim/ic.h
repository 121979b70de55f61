// ic.h -- initial-condition factory (src/ic/ic.h:9-18, src/ic/solitons.h)
#ifndef QX_SHIM_IC_H
#define QX_SHIM_IC_H

#include <memory>
#include <string>
#include "field.h"

class InitialCondition {
  public:
    virtual ~InitialCondition() = default;
    virtual void apply(Field &field) const = 0;
    virtual std::string name() const = 0;
};

// Gaussian solitons from the table ./solitons.txt, formed on the device
// (src/ic/solitons.cxx:20-110)
class SolitonsIC : public InitialCondition {
  public:
    explicit SolitonsIC(const Params &p) : p_(p) {}
    void apply(Field &field) const override;
    std::string name() const override { return "solitons"; }

  private:
    const Params &p_;
};

std::unique_ptr<InitialCondition> createIC(ICType type, const Params &p);

#endif
