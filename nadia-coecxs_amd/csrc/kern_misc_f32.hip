#define REAL float
#include "kern_misc.inc"
