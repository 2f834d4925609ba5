/* Compiled into oracle/_ref/libsp3mm_ref.so with the reference's own headers
 * (TEST INFRASTRUCTURE ONLY, never part of the product).
 *
 * The reference's test driver installs the reference's OpenMP chunk
 * distribution function as CONFIG.chunkDistrbFunc
 * (code/test/Sp3MM_test.c:80, :304-308): chunksFairFolded
 * (code/src/include/ompChunksDivide.h:77), an inline function of that header.
 * The declaration below makes this unit emit its external definition, and the
 * getter hands its address to the CPU baseline, so that the reference runs
 * under the same OpenMP schedule adaptation as in its own driver. */
#include <omp.h>
#include "macros.h"
#include "sparseMatrix.h"
#include "utils.h"
#include "ompChunksDivide.h"

void chunksFairFolded(ulong r, spmat *mat, CONFIG *cfg);

void *sp3mm_ref_chunks_fair_folded(void) { return (void *)&chunksFairFolded; }
