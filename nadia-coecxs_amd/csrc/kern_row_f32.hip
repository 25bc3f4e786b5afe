#define REAL float
#define NADIA_PASS_ROW
#include "kern_pass.inc"
