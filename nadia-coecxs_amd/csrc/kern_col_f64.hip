#define REAL double
#define NADIA_PASS_COL
#include "kern_pass.inc"
