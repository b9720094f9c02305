/* Oracle variant: multi, 2-D.  Test infrastructure only (see oracle_common.h). */
#define OD 2
#define OMODEL O_MULTI
#define OPFX(n) o_multi2_##n
#include "oracle_impl.h"
