/* TEST INFRASTRUCTURE ONLY: the real_t = float instantiation of the oracle. */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "oracle.h"

#define R float
#define S _f
#include "oracle_impl.inc"
#include "oracle_deck.inc"
#include "oracle_adjoint.inc"
