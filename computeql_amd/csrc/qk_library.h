/*
 * qk_library.h - registry of the library kernels compiled into the backend
 * (the HIP counterpart of the named .comp shaders that gpu_kernel_alloc
 * resolves through the SPIR-V provider, query_kernel.c:459-494).
 */
#ifndef QK_LIBRARY_H
#define QK_LIBRARY_H

#include "../../include/cqlhip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct QK_LibraryKernel {
  const char* name;
  const void* host_fn;  /* __global__ void fn(QK_Args) */
  U32 block;            /* fixed workgroup size */
} QK_LibraryKernel;

const QK_LibraryKernel* qk_library_find(const char* name);
U32 qk_library_count(void);
const QK_LibraryKernel* qk_library_at(U32 index);

#ifdef __cplusplus
}
#endif
#endif
