#define REAL double
#include "kern_pass.inc"
