/* Oracle variant: euler, 2-D.  Test infrastructure only (see oracle_common.h). */
#define OD 2
#define OMODEL O_EULER
#define OPFX(n) o_euler2_##n
#include "oracle_impl.h"
