// isti_internal.hpp - entry points shared between the library's own source
// files (not part of the C ABI).
#ifndef ISTI_INTERNAL_HPP
#define ISTI_INTERNAL_HPP

#include <stddef.h>
#include <stdint.h>

#include <string>

#include "../../include/isti.h"

// Scratch bytes a frame of `info` needs when all its scratch planes are
// carved from one block of device memory (each plane 256-byte aligned).
size_t isti_internal_scratch_bytes(const isti_renderer *r, const isti_render_info *info);

// isti_render with the renderer's own scratch (ext_base null), or with every
// scratch plane carved from ext_bytes of device memory at ext_base: then
// nothing is allocated, frames needing more than ext_bytes are rejected with
// ISTI_ERR_UNSUPPORTED_SHAPE, and every plane is rewritten before it is read.
isti_status isti_internal_render(isti_renderer *r, const isti_render_info *info, uint8_t *out, size_t out_pitch,
                                 void *stream, void *ext_base, size_t ext_bytes);

// Device index the renderer was created on.
int isti_internal_device(const isti_renderer *r);

// Records msg as the renderer's last error and returns st.
isti_status isti_internal_fail(isti_renderer *r, isti_status st, const std::string &msg);

#endif
