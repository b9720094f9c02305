/* Oracle variant: ns, 3-D.  Test infrastructure only (see oracle_common.h). */
#define OD 3
#define OMODEL O_NAVIER_STOKES
#define OPFX(n) o_ns3_##n
#include "oracle_impl.h"
