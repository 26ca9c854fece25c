// One-launch adjoint step blocks (joint.h), float and double, depths 1, 2, 4.
#include "joint.h"

namespace etl {

template <typename R>
int launch_joint_adj_any(const AdjSet<R>* a, int ta, hipStream_t s) {
    switch (ta) {
        case 1: return launch_joint_t<R, 0, 1>(nullptr, a, s);
        case 2: return launch_joint_t<R, 0, 2>(nullptr, a, s);
        case 4: return launch_joint_t<R, 0, 4>(nullptr, a, s);
    }
    set_error("joint step: adjoint depth %d not compiled", ta);
    return ETL_EARG;
}

template <>
int launch_joint_adj<float>(const AdjSet<float>* a, int ta, hipStream_t s) {
    return launch_joint_adj_any<float>(a, ta, s);
}
template <>
int launch_joint_adj<double>(const AdjSet<double>* a, int ta, hipStream_t s) {
    return launch_joint_adj_any<double>(a, ta, s);
}

}  // namespace etl
