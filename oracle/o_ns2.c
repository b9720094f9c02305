/* Oracle variant: ns, 2-D.  Test infrastructure only (see oracle_common.h). */
#define OD 2
#define OMODEL O_NAVIER_STOKES
#define OPFX(n) o_ns2_##n
#include "oracle_impl.h"
