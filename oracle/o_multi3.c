/* Oracle variant: multi, 3-D.  Test infrastructure only (see oracle_common.h). */
#define OD 3
#define OMODEL O_MULTI
#define OPFX(n) o_multi3_##n
#include "oracle_impl.h"
