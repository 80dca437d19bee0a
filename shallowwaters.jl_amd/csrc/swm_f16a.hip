// swm_f16a.hip -- the step and ensemble launches for the number-format bundle FmtF16A
// (Float16 values with float arithmetic, Tprog = Float32), in a translation unit of their own.
#include "swm_launch.h"

const FmtOps &ops_f16a() {
    static const FmtOps o = Entry<FmtF16A, false>::make();
    return o;
}
