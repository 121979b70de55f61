// enum.h -- enumerations of the reference's configuration (src/enum.h:6-47),
// kept value-for-value so Params built for the reference mean the same here.
#ifndef QX_SHIM_ENUM_H
#define QX_SHIM_ENUM_H

#include <cstdint>

enum class ICType { SOLITONS, SPECTRUM };
enum class CosmoType { STATIC, MRE };
// FFTW planning rigour (src/enum.h:18-24).  The GPU transforms have one
// fixed schedule per N, so every value selects the same plan.
enum class FFTPlanType { ESTIMATE, MEASURE, PATIENT, EXHAUSTIVE };
enum class MeasureType : uint32_t {
    NONE = 0,
    SPECTRUM = 1 << 0,
    RHO_MAX = 1 << 1,
    RHO_SLICE = 1 << 2,
    RHO_GRID = 1 << 3,
    PSI_GRID = 1 << 4
};
// arithmetic of the device state (no reference counterpart: the reference is FP64)
enum class Precision { FP64 = 1, FP32 = 0 };

inline MeasureType operator|(MeasureType a, MeasureType b)
{
    return static_cast<MeasureType>(static_cast<uint32_t>(a) | static_cast<uint32_t>(b));
}
inline bool operator&(MeasureType a, MeasureType b)
{
    return (static_cast<uint32_t>(a) & static_cast<uint32_t>(b)) != 0;
}

#endif
