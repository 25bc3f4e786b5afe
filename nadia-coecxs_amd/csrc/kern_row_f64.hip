#define REAL double
#define NADIA_PASS_ROW
#include "kern_pass.inc"
