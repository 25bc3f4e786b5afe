#define REAL double
#include "kern_misc.inc"
