// One-launch step blocks of a forward and an adjoint plan together (joint.h),
// float, depths 1, 2, 4, 6, forward interior tiles of 8 or 16 rows.
#include "joint.h"

namespace etl {

template <>
int launch_joint_pair<float>(const FwdSet<float>* f, const AdjSet<float>* a, int t, int part,
                            hipStream_t s) {
    if (f->band == 8) {
        switch (t) {
            case 1: return launch_joint_part<float, 1, 1>(f, a, part, s);
            case 2: return launch_joint_part<float, 2, 2>(f, a, part, s);
            case 4: return launch_joint_part<float, 4, 4>(f, a, part, s);
            case 6: return launch_joint_part<float, 6, 6>(f, a, part, s);
        }
    } else if (f->band == 16) {
        switch (t) {
            case 1: return launch_joint_part<float, 1, 1, 16>(f, a, part, s);
            case 2: return launch_joint_part<float, 2, 2, 16>(f, a, part, s);
            case 4: return launch_joint_part<float, 4, 4, 16>(f, a, part, s);
            case 6: return launch_joint_part<float, 6, 6, 16>(f, a, part, s);
        }
    }
    set_error("joint step: depth %d, forward band %d not compiled", t, f->band);
    return ETL_EARG;
}

int warm_scratch_pair(hipStream_t s) {
    return launch_joint_empty<float, 6, 6, kAllKinds, 8>(kWarmGroups, s);
}

}  // namespace etl
