/* Oracle variant: euler, 3-D.  Test infrastructure only (see oracle_common.h). */
#define OD 3
#define OMODEL O_EULER
#define OPFX(n) o_euler3_##n
#include "oracle_impl.h"
