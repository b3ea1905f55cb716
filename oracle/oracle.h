/* oracle.h -- CPU restatement of the reference path (test infrastructure only). */
#ifndef BTC_ORACLE_H
#define BTC_ORACLE_H
#include <stdint.h>

int64_t orc_canon(const int32_t* src, const int32_t* dst, int64_t nnz, int32_t n, uint32_t flags,
                  int64_t** rowptr_out, int32_t** col_out, int32_t** labels_out, int32_t* n_out);
void orc_free(void* p);
uint64_t orc_tc_merge(const int64_t* rowptr, const int32_t* col, int32_t n, int32_t threads, int64_t* per_row);
uint64_t orc_tc_brute(const int64_t* rowptr, const int32_t* col, int32_t n, int64_t* per_row);
int64_t orc_bitbsr_build(const int64_t* rowptr, const int32_t* col, int32_t n, int32_t K, uint32_t* indptr,
                         uint32_t* indices, uint32_t* row_indices, uint64_t* tiles);
uint64_t orc_bitbsr_count(int32_t K, int32_t n, const uint32_t* indptr, const uint32_t* indices,
                          const uint32_t* row_indices, const uint64_t* tiles, uint32_t nblk, int32_t tau,
                          int32_t variant, int64_t* per_row, int64_t* dense_tiles);
int32_t orc_lite_tile_k(int32_t n, int64_t m, int32_t mode);
int32_t orc_variant(int32_t K, uint32_t nblk, int32_t threshold_blocks);
#endif
