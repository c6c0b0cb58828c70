/* TEST INFRASTRUCTURE ONLY: the real_t = double oracle with the sqrtf
 * narrowing of push.h:148,169 replaced by a double sqrt (suffix _ds).
 *
 * The reference's double build takes gamma through sqrtf, so its step is a
 * staircase at float resolution (steps of about 6e-8 relative in gamma) and
 * a central difference of it with a small epsilon measures the staircase,
 * not the slope.  This smooth instantiation of the same source is the
 * finite-difference reference for the tangent-linear step; it is never
 * compared bit for bit with anything. */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "oracle.h"

#define R double
#define S _ds
#define sqrtf sqrt
#include "oracle_impl.inc"
#include "oracle_deck.inc"
#include "oracle_adjoint.inc"
