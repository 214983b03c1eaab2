// Element restriction E (gather) and a bit-exact E^T (ordered gather-sum, no atomics).
//
// E-vector layout is element-major [elem][comp][node], the layout of the
// reference CPU backend (backends/ref/ceed-ref-restriction.c:768-772), so the
// E-vectors of this backend and of /cpu/self/ref/serial are byte-comparable.
//
// E^T is computed per L-vector entry from a transpose map built on the host:
// for entry t, the E-vector positions that map to t, in ascending
// (elem, comp, node) order.  One thread owns t, starts from the current v[t]
// and adds the contributions in that order, which is exactly the sequence of
// roundings of the reference's serial loop (ceed-ref-restriction.c:220-242).
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "mi355_common.h"

struct CeedMi355Restriction_private {
  int32_t  num_elem, elem_size, num_comp, comp_stride;
  int64_t  l_size;
  bool     strided;
  int32_t  strides[3];
  int32_t *d_offsets;    // [num_elem * elem_size] (offset restriction)
  // transpose map over the L entries that receive contributions
  int64_t  num_targets;
  int64_t *d_t_list;     // L index of each target
  int64_t *d_t_start;    // [num_targets + 1] into d_t_index
  int64_t *d_t_index;    // E-vector positions, ascending (elem, comp, node)
};

namespace {

__global__ void __launch_bounds__(256) offset_notranspose(int64_t n, int32_t elem_size, int32_t num_comp, int32_t comp_stride,
                                                          const int32_t *__restrict__ offsets, const double *__restrict__ u,
                                                          double *__restrict__ v) {
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < n; idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t e = idx / elem_size, i = idx - e * elem_size;
    const int64_t o = offsets[idx];
    for (int32_t k = 0; k < num_comp; k++) v[(e * num_comp + k) * elem_size + i] = u[o + (int64_t)k * comp_stride];
  }
}

__global__ void __launch_bounds__(256) strided_notranspose(int64_t n, int32_t elem_size, int32_t num_comp, int32_t s0, int32_t s1,
                                                           int32_t s2, const double *__restrict__ u, double *__restrict__ v) {
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < n; idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t per = (int64_t)elem_size * num_comp;
    const int64_t e = idx / per, r = idx - e * per, k = r / elem_size, i = r - k * elem_size;
    v[idx] = u[i * s0 + k * s1 + e * s2];
  }
}

__global__ void __launch_bounds__(256) ordered_transpose(int64_t num_targets, const int64_t *__restrict__ t_list,
                                                         const int64_t *__restrict__ t_start, const int64_t *__restrict__ t_index,
                                                         const double *__restrict__ u, double *__restrict__ v) {
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < num_targets; j += (int64_t)gridDim.x * blockDim.x) {
    const int64_t t = t_list[j];
    double        acc = v[t];
    for (int64_t m = t_start[j]; m < t_start[j + 1]; m++) acc += u[t_index[m]];
    v[t] = acc;
  }
}

// Build the transpose map from the L index of every E-vector position.
int build_transpose_map(CeedMi355Restriction r, const std::vector<int64_t> &l_of_e) {
  const int64_t         n_e = (int64_t)l_of_e.size();
  std::vector<int64_t>  count(r->l_size + 1, 0);
  for (int64_t j = 0; j < n_e; j++) {
    MI355_CHECK(l_of_e[j] >= 0 && l_of_e[j] < r->l_size, CEED_MI355_ERROR_DIMENSION,
                "restriction maps E-vector entry %lld to L index %lld outside [0, %lld)", (long long)j, (long long)l_of_e[j],
                (long long)r->l_size);
    count[l_of_e[j]]++;
  }
  std::vector<int64_t> list, start;
  std::vector<int64_t> slot(r->l_size, -1);
  for (int64_t t = 0; t < r->l_size; t++)
    if (count[t]) {
      slot[t] = (int64_t)list.size();
      list.push_back(t);
    }
  start.assign(list.size() + 1, 0);
  for (size_t j = 0; j < list.size(); j++) start[j + 1] = start[j] + count[list[j]];
  std::vector<int64_t> cursor(start.begin(), start.end() - 1), index(n_e);
  // E-vector positions are visited in ascending (elem, comp, node) order
  for (int64_t e = 0; e < r->num_elem; e++)
    for (int64_t k = 0; k < r->num_comp; k++)
      for (int64_t i = 0; i < r->elem_size; i++) {
        const int64_t epos = (e * r->num_comp + k) * r->elem_size + i;
        const int64_t t    = l_of_e[epos];
        index[cursor[slot[t]]++] = epos;
      }
  r->num_targets = (int64_t)list.size();
  MI355_HIP(hipMalloc(&r->d_t_list, sizeof(int64_t) * (list.size() + 1)));
  MI355_HIP(hipMalloc(&r->d_t_start, sizeof(int64_t) * start.size()));
  MI355_HIP(hipMalloc(&r->d_t_index, sizeof(int64_t) * (index.size() + 1)));
  MI355_HIP(hipMemcpy(r->d_t_list, list.data(), sizeof(int64_t) * list.size(), hipMemcpyHostToDevice));
  MI355_HIP(hipMemcpy(r->d_t_start, start.data(), sizeof(int64_t) * start.size(), hipMemcpyHostToDevice));
  MI355_HIP(hipMemcpy(r->d_t_index, index.data(), sizeof(int64_t) * index.size(), hipMemcpyHostToDevice));
  return CEED_MI355_SUCCESS;
}

CeedMi355Restriction alloc_rstr(int32_t num_elem, int32_t elem_size, int32_t num_comp, int64_t l_size) {
  CeedMi355Restriction r = (CeedMi355Restriction)calloc(1, sizeof(*r));
  r->num_elem = num_elem;
  r->elem_size = elem_size;
  r->num_comp = num_comp;
  r->l_size = l_size;
  return r;
}

}  // namespace

extern "C" int CeedMi355RestrictionCreateOffset(int32_t num_elem, int32_t elem_size, int32_t num_comp, int32_t comp_stride, int64_t l_size,
                                                const int32_t *offsets_host, CeedMi355Restriction *rstr) {
  *rstr = nullptr;
  MI355_CHECK(num_elem >= 0 && elem_size > 0 && num_comp > 0 && l_size >= 0, CEED_MI355_ERROR_DIMENSION, "invalid restriction sizes");
  CeedMi355Restriction r = alloc_rstr(num_elem, elem_size, num_comp, l_size);
  r->comp_stride = comp_stride;
  const int64_t n = (int64_t)num_elem * elem_size;
  std::vector<int64_t> l_of_e((size_t)(n * num_comp));
  for (int64_t e = 0; e < num_elem; e++)
    for (int64_t k = 0; k < num_comp; k++)
      for (int64_t i = 0; i < elem_size; i++)
        l_of_e[(e * num_comp + k) * elem_size + i] = (int64_t)offsets_host[e * elem_size + i] + k * (int64_t)comp_stride;
  int ierr = build_transpose_map(r, l_of_e);
  if (ierr) {
    CeedMi355RestrictionDestroy(&r);
    return ierr;
  }
  MI355_HIP(hipMalloc(&r->d_offsets, sizeof(int32_t) * (n + 1)));
  MI355_HIP(hipMemcpy(r->d_offsets, offsets_host, sizeof(int32_t) * n, hipMemcpyHostToDevice));
  *rstr = r;
  return CEED_MI355_SUCCESS;
}

extern "C" int CeedMi355RestrictionCreateStrided(int32_t num_elem, int32_t elem_size, int32_t num_comp, int64_t l_size,
                                                 const int32_t *strides_host, CeedMi355Restriction *rstr) {
  *rstr = nullptr;
  MI355_CHECK(num_elem >= 0 && elem_size > 0 && num_comp > 0 && l_size >= 0, CEED_MI355_ERROR_DIMENSION, "invalid restriction sizes");
  CeedMi355Restriction r = alloc_rstr(num_elem, elem_size, num_comp, l_size);
  r->strided = true;
  memcpy(r->strides, strides_host, sizeof(r->strides));
  std::vector<int64_t> l_of_e((size_t)num_elem * num_comp * elem_size);
  for (int64_t e = 0; e < num_elem; e++)
    for (int64_t k = 0; k < num_comp; k++)
      for (int64_t i = 0; i < elem_size; i++)
        l_of_e[(e * num_comp + k) * elem_size + i] = i * r->strides[0] + k * r->strides[1] + e * r->strides[2];
  int ierr = build_transpose_map(r, l_of_e);
  if (ierr) {
    CeedMi355RestrictionDestroy(&r);
    return ierr;
  }
  *rstr = r;
  return CEED_MI355_SUCCESS;
}

extern "C" int CeedMi355RestrictionApply(CeedMi355Restriction r, int t_mode, const double *u, double *v, void *stream) {
  MI355_CHECK(r, CEED_MI355_ERROR_MAJOR, "null restriction");
  hipStream_t   s = (hipStream_t)stream;
  const int64_t n = (int64_t)r->num_elem * r->elem_size;
  if (n == 0) return CEED_MI355_SUCCESS;
  if (t_mode == 0) {
    if (r->strided) {
      strided_notranspose<<<mi355_blocks(n * r->num_comp, 256), 256, 0, s>>>(n * r->num_comp, r->elem_size, r->num_comp, r->strides[0],
                                                                              r->strides[1], r->strides[2], u, v);
    } else {
      offset_notranspose<<<mi355_blocks(n, 256), 256, 0, s>>>(n, r->elem_size, r->num_comp, r->comp_stride, r->d_offsets, u, v);
    }
  } else {
    if (r->num_targets == 0) return CEED_MI355_SUCCESS;
    ordered_transpose<<<mi355_blocks(r->num_targets, 256), 256, 0, s>>>(r->num_targets, r->d_t_list, r->d_t_start, r->d_t_index, u, v);
  }
  MI355_LAUNCH_CHECK();
  return CEED_MI355_SUCCESS;
}

extern "C" int CeedMi355RestrictionDestroy(CeedMi355Restriction *rstr) {
  if (!rstr || !*rstr) return CEED_MI355_SUCCESS;
  CeedMi355Restriction r = *rstr;
  (void)hipFree(r->d_offsets);
  (void)hipFree(r->d_t_list);
  (void)hipFree(r->d_t_start);
  (void)hipFree(r->d_t_index);
  free(r);
  *rstr = nullptr;
  return CEED_MI355_SUCCESS;
}
