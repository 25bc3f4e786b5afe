#define REAL float
#include "kern_pass.inc"
