// One-launch forward step blocks (joint.h), float and double, depths 1, 2, 4, 6, and 8
// in float with 16-row tiles.
#include "joint.h"

namespace etl {

template <typename R>
int launch_joint_fwd_any(const FwdSet<R>* f, int tf, int part, hipStream_t s) {
    if (f->band == 8) {
        switch (tf) {
            case 1: return launch_joint_part<R, 1, 0>(f, nullptr, part, s);
            case 2: return launch_joint_part<R, 2, 0>(f, nullptr, part, s);
            case 4: return launch_joint_part<R, 4, 0>(f, nullptr, part, s);
        }
        if (tf == 6) return launch_joint_part<R, 6, 0>(f, nullptr, part, s);
    } else if (f->band == 16) {
        switch (tf) {
            case 1: return launch_joint_part<R, 1, 0, 16>(f, nullptr, part, s);
            case 2: return launch_joint_part<R, 2, 0, 16>(f, nullptr, part, s);
            case 4: return launch_joint_part<R, 4, 0, 16>(f, nullptr, part, s);
            case 6: return launch_joint_part<R, 6, 0, 16>(f, nullptr, part, s);
        }
        if constexpr (sizeof(R) == 4)
            if (tf == 8) return launch_joint_part<R, 8, 0, 16>(f, nullptr, part, s);
    }
    set_error("joint step: forward depth %d, band %d not compiled", tf, f->band);
    return ETL_EARG;
}

template <>
int launch_joint_fwd<float>(const FwdSet<float>* f, int tf, int part, hipStream_t s) {
    return launch_joint_fwd_any<float>(f, tf, part, s);
}
template <>
int launch_joint_fwd<double>(const FwdSet<double>* f, int tf, int part, hipStream_t s) {
    return launch_joint_fwd_any<double>(f, tf, part, s);
}

template <typename R>
int launch_joint_fwd_empty_any(int tf, int band, hipStream_t s) {
    if (band == 8) {
        switch (tf) {
            case 1: return launch_joint_empty<R, 1, 0, kAllKinds, 8>(1, s);
            case 2: return launch_joint_empty<R, 2, 0, kAllKinds, 8>(1, s);
            case 4: return launch_joint_empty<R, 4, 0, kAllKinds, 8>(1, s);
            case 6: return launch_joint_empty<R, 6, 0, kAllKinds, 8>(1, s);
        }
    } else if (band == 16) {
        switch (tf) {
            case 1: return launch_joint_empty<R, 1, 0, kAllKinds, 16>(1, s);
            case 2: return launch_joint_empty<R, 2, 0, kAllKinds, 16>(1, s);
            case 4: return launch_joint_empty<R, 4, 0, kAllKinds, 16>(1, s);
            case 6: return launch_joint_empty<R, 6, 0, kAllKinds, 16>(1, s);
        }
        if constexpr (sizeof(R) == 4)
            if (tf == 8) return launch_joint_empty<R, 8, 0, kAllKinds, 16>(1, s);
    }
    set_error("joint step: forward depth %d, band %d not compiled", tf, band);
    return ETL_EARG;
}

template <>
int launch_joint_fwd_empty<float>(int tf, int band, hipStream_t s) {
    return launch_joint_fwd_empty_any<float>(tf, band, s);
}
template <>
int launch_joint_fwd_empty<double>(int tf, int band, hipStream_t s) {
    return launch_joint_fwd_empty_any<double>(tf, band, s);
}

int warm_scratch_fwd(hipStream_t s) {
    return launch_joint_empty<double, 6, 0, kAllKinds, 8>(kWarmGroups, s);
}

int warm_scratch(hipStream_t s) {
    if (int rc = warm_scratch_fwd(s)) return rc;
    return warm_scratch_pair(s);
}

}  // namespace etl
