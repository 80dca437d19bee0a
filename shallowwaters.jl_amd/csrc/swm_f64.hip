// swm_f64.hip -- the step, ensemble and reverse-mode launches for the number-format
// bundle FmtF64, in a translation unit of their own (the units compile in parallel).
#include "swm_launch.h"

const FmtOps &ops_f64() {
    static const FmtOps o = Entry<FmtF64, true>::make();
    return o;
}
