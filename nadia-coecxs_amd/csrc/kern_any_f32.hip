#define REAL float
#include "kern_any.inc"
