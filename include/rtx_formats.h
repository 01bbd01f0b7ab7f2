/*
 * rtx_formats.h — the reference's text and on-disk formats around the
 * traversal path, for interchange with its tooling.  Host-only.
 */
#ifndef RTX_FORMATS_H
#define RTX_FORMATS_H

#include <stddef.h>
#include <stdint.h>

#include "rtx.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Per-ray treelet trace printed by traceRayWithTreelets
 * (src/cuda-sim/vulkan_ray_tracing.cc:2256-2262): one line per ray,
 *   RayID,<id>,0x<treelet>,0x<treelet>,...\n
 * with one entry per memory transaction (leaf-descriptor records included)
 * and each treelet named by its root address printed with %x (low 32 bits).
 * Here the root address is base_address + the root's arena offset
 * (roots[record.treelet]); ray i gets id first_ray_id + i (the reference counts
 * from 1 in issue order).  records/offsets are the CSR output of rtx_trace.
 * Writes at most `capacity` bytes to `out` (NUL-terminated when room is left)
 * and the full length to *length; out = NULL only measures.  A record without
 * a treelet returns RTX_ERR_NO_PARTITION (the reference asserts, :470). */
rtx_status rtx_format_treelet_trace(const rtx_record* records, const uint64_t* offsets, uint32_t n_rays,
                                    const uint64_t* roots, uint32_t n_treelets, uint64_t base_address,
                                    uint32_t first_ray_id, char* out, size_t capacity, size_t* length);

/* Split-file acceleration-structure dump of dump_descriptor_set_for_AS
 * (src/cuda-sim/vulkan_ray_tracing.cc:4455-4558, split_files form) with
 * findOffsetBounds (:4901-4944):
 *   <prefix>.asmain      the top-level structure (desc_size bytes from the TLAS)
 *   <prefix>.asback      bottom levels below the TLAS (when there are any)
 *   <prefix>.asfront     bottom levels above the TLAS (when there are any)
 *   <prefix>.asmetadata  "desc_size,1000150000,max_backwards,min_backwards,
 *                         min_forwards,max_forwards,back_buffer,front_buffer,
 *                         haveBackwards,haveForwards"
 * Offsets are BLAS header offsets relative to the TLAS header.  The writer
 * sets the two buffer amounts to the extent of the last bottom level on each
 * side, so that the files hold every byte of every BLAS. */
rtx_status rtx_as_dump_write(const void* arena_host, uint64_t arena_bytes, uint64_t tlas_offset, const char* prefix);
/* Arena size and TLAS offset that rtx_as_dump_read will produce for a dump. */
rtx_status rtx_as_dump_info(const char* prefix, uint64_t* arena_bytes, uint64_t* tlas_offset);
/* Rebuild one arena from a dump: every region at its offset from the TLAS,
 * the gaps zero-filled, the TLAS at a 64-byte aligned offset. */
rtx_status rtx_as_dump_read(const char* prefix, void* arena_host, uint64_t capacity);

#ifdef __cplusplus
}
#endif
#endif /* RTX_FORMATS_H */
