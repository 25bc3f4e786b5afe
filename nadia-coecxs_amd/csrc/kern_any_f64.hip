#define REAL double
#include "kern_any.inc"
