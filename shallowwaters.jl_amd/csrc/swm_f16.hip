// swm_f16.hip -- the step and ensemble launches for the number-format
// bundle FmtF16, in a translation unit of their own (the units compile in parallel).
#include "swm_launch.h"

const FmtOps &ops_f16() {
    static const FmtOps o = Entry<FmtF16, false>::make();
    return o;
}
