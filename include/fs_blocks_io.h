/*
 * fs_blocks_io.h -- binary interchange of an assembled Stokes system.
 *
 * The reference exports its assembled matrix at the EXPORT_MATRIX hook of
 * StokesSimulatorBase::advance_time (include/felspa/pde/src/
 * stokes_common.implement.h:283-287) as formatted text. This format carries
 * what the solve needs, in binary: the blocks A, B^T, B, K11 of
 * K = [[A, B^T], [B, K11]] and M_p = preconditioner(1,1), each as CSR in the
 * order the writer stores it (deal.II: diagonal first in square blocks, the
 * flag kept per block), and optionally the right-hand side. A deal.II program
 * writes it with fs_blocks_save from its SparseMatrix iterators; the loaded
 * arrays feed fs_stokes_create directly.
 *
 * File layout (little-endian): "FSBK", uint32 version (1), int64 n_u, int64
 * n_p, uint32 has_rhs, uint32 0; per block: int64 n_rows, n_cols, nnz,
 * int32 diag_first, int32 0, int64 row_ptr[n_rows+1], int32 col_idx[nnz],
 * double val[nnz]; double rhs[n_u+n_p] when has_rhs; uint64 FNV-1a hash of
 * every byte after the header. Host only: no device is needed.
 */
#ifndef FS_BLOCKS_IO_H
#define FS_BLOCKS_IO_H

#include <stdint.h>

#include "fs_stokes.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Writes the five blocks (A [n_u x n_u], Bt [n_u x n_p], B [n_p x n_u],
 * K11 [n_p x n_p], Mp [n_p x n_p]) and rhs [n_u + n_p] (may be NULL). */
int fs_blocks_save(const char* path, const fs_csr* A, const fs_csr* Bt, const fs_csr* B, const fs_csr* K11,
                   const fs_csr* Mp, const double* rhs);

typedef struct fs_blocks fs_blocks;
/* Reads and checks a file (sizes, index ranges, hash); FS_ERR_BAD_INPUT on a
 * malformed or corrupted file. */
int fs_blocks_load(const char* path, fs_blocks** out);
/* Borrowed view of block `which` (0 = A, 1 = Bt, 2 = B, 3 = K11, 4 = Mp),
 * valid until fs_blocks_free. */
int fs_blocks_get(const fs_blocks* f, int32_t which, fs_csr* out);
/* rhs [n_u + n_p] or NULL when the file has none; *n_u / *n_p if given. */
const double* fs_blocks_rhs(const fs_blocks* f, int64_t* n_u, int64_t* n_p);
void fs_blocks_free(fs_blocks* f);

#ifdef __cplusplus
}
#endif
#endif
