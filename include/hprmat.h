/*
 * hprmat.h - the header name of HPRMAT's C binding (bindings/c/hprmat.h).
 *
 * A C program written against the reference's binding includes "hprmat.h";
 * with this directory on its include path it compiles unchanged against
 * libhprmat_hip.  Every declaration (hprmat_init, hprmat_solve, ...,
 * hprmat_gpu_finalize and the HPRMAT_SOLVER_* constants) lives in
 * hprmat_hip.h, Part 3.  Complex arrays are declared void * there, so the
 * reference's double _Complex * arguments convert implicitly.
 */
#ifndef HPRMAT_H
#define HPRMAT_H
#include "hprmat_hip.h"
#endif
