// One-launch adjoint step blocks (joint.h), float and double, depths 1, 2, 4 (and 6 in
// float).
#include "joint.h"

namespace etl {

template <typename R>
int launch_joint_adj_any(const AdjSet<R>* a, int ta, int part, hipStream_t s) {
    switch (ta) {
        case 1: return launch_joint_part<R, 0, 1>(nullptr, a, part, s);
        case 2: return launch_joint_part<R, 0, 2>(nullptr, a, part, s);
        case 4: return launch_joint_part<R, 0, 4>(nullptr, a, part, s);
    }
    if constexpr (sizeof(R) == 4) {  // depth 6: float only
        if (ta == 6) return launch_joint_part<R, 0, 6>(nullptr, a, part, s);
    }
    set_error("joint step: adjoint depth %d not compiled", ta);
    return ETL_EARG;
}

template <>
int launch_joint_adj<float>(const AdjSet<float>* a, int ta, int part, hipStream_t s) {
    return launch_joint_adj_any<float>(a, ta, part, s);
}
template <>
int launch_joint_adj<double>(const AdjSet<double>* a, int ta, int part, hipStream_t s) {
    return launch_joint_adj_any<double>(a, ta, part, s);
}

}  // namespace etl
