/* aqflip_hash.h -- spatial hashing service of libaqflip (C ABI).
 *
 * The reference sorts particles by grid cell with performSpatialHashing2
 * (Fluid/Fluid_3D_common.cuh:28-56), shared by FLIP (Fluid/Fluid_3D_FLIP.cu:364),
 * PBF (Fluid/Fluid_3D_PBF.cu:474,483,491), PCISPH (Fluid/Fluid_3D_PCISPH.cu:420,434)
 * and the mesher's SDF grid (Rendering/Renderer3D/Mesher.cuh:155). This is the same
 * operation as a standalone handle: the cell hash of calcHashImpl
 * (Fluid/Fluid_3D_kernels.cuh:8-27), a stable sort by hash (sort_by_key with the
 * identity indices, Fluid/Fluid_3D_common.cuh:31-35; stable by input index), the
 * reordered records (applySortImpl, Fluid/Fluid_3D_kernels.cuh:95-102) and the cell
 * ranges (findCellStartEndImpl, Fluid/Fluid_kernels.cuh:3-22).
 *
 * hash = x*(size_y*size_z) + y*size_z + z with x = (int)(px / cell_size) (truncating,
 * IEEE division), clamped to the grid on every axis (the reference indexes out of
 * bounds there). Records are `stride` floats whose first three are the position:
 * stride 3 for packed positions, 12 for the reference's 48-byte Particle.
 *
 * Threading: one handle per thread; calls are synchronous at return. Calls on one
 * handle must not overlap (they share the handle's bins).
 */
#ifndef AQFLIP_HASH_H
#define AQFLIP_HASH_H

#include <stddef.h>
#include <stdint.h>

#include "aqflip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct aq_hash aq_hash;

/* Grid of size_x*size_y*size_z cells of edge cell_size. capacity: expected particle
 * count (the handle grows on demand). */
aq_status aq_hash_create(int32_t size_x, int32_t size_y, int32_t size_z, float cell_size,
                         int64_t capacity, aq_hash** out);
void aq_hash_destroy(aq_hash* h);

/* Host arrays. records: n records of `stride` floats. Outputs (each may be NULL):
 *   perm[n]             sorted slot -> input index (the sorted particleIndices)
 *   sorted_hash[n]      the sorted particleHashes
 *   cell_start[cell_count], cell_end[cell_count]  first / last (inclusive) sorted slot
 *                       of each cell, -1 when the cell is empty. cell_count is the
 *                       length of these arrays: 0 means size_x*size_y*size_z; a larger
 *                       count sets the entries past the grid's cells to -1, as the
 *                       reference's clear of cellCount entries does
 *                       (Fluid/Fluid_3D_common.cuh:47-48; the mesher passes an SDF cell
 *                       count above the hashed grid's, Rendering/Renderer3D/Mesher.cuh:155)
 *   sorted_records[n*stride]  the records in sorted order
 * Positions must be finite. */
aq_status aq_hash_sort_host(aq_hash* h, const float* records, int64_t n, int32_t stride,
                            int32_t* perm, int32_t* sorted_hash, int32_t* cell_start,
                            int32_t* cell_end, int64_t cell_count, float* sorted_records);

/* The same on device memory allocated by the caller (hipMalloc). stream is the
 * caller's hipStream_t (NULL: the null stream): the kernels are queued on it, so they
 * run after the work the caller queued there before the call (for instance the
 * kernel that wrote the records), and the call returns once they are complete (it
 * synchronises that stream, as the reference synchronises the device after its sort,
 * Fluid/Fluid_3D_common.cuh:38). Work on other streams is not waited for. A
 * non-finite position hashes to a clamped cell instead of being rejected. */
aq_status aq_hash_sort_device(aq_hash* h, const float* d_records, int64_t n, int32_t stride,
                              int32_t* d_perm, int32_t* d_sorted_hash, int32_t* d_cell_start,
                              int32_t* d_cell_end, int64_t cell_count, float* d_sorted_records,
                              void* stream);

/* No counterpart in the reference. For callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): aq_hash_sort_device without a handle, calling no hipMalloc/hipFree.
 * Its bins (per-cell counts and offsets, per-record hash, rank and permutation) live
 * in one static device array of 69 MB that every GPU the library runs on holds until
 * the process ends, so it takes at most 2,097,152 cells (size_x*size_y*size_z) and
 * 4,194,304 records; larger requests fail with AQ_EINVAL. Inputs, outputs, cell_count
 * and stream ordering are those of aq_hash_sort_device. Every call on one GPU shares
 * that GPU's array: calls that overlap on one GPU (other host threads, other streams)
 * race and must be serialised by the caller. */
aq_status aq_hash_sort_device_noalloc(int32_t size_x, int32_t size_y, int32_t size_z, float cell_size,
                                      const float* d_records, int64_t n, int32_t stride,
                                      int32_t* d_perm, int32_t* d_sorted_hash, int32_t* d_cell_start,
                                      int32_t* d_cell_end, int64_t cell_count,
                                      float* d_sorted_records, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* AQFLIP_HASH_H */
