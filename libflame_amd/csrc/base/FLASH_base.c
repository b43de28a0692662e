/*
 * FLASH_base.c -- hierarchical (storage-by-blocks) objects of libflame_amd.
 *
 * Follows (paths relative to the libflame source tree):
 *   FLASH_Obj_create{,_ext,_without_buffer}  src/base/flamec/hierarchy/main/FLASH_Obj.c:143-269
 *   leaf storage order                       FLASH_Obj.c:271-404 (FLASH_Obj_create_hierarchy)
 *   conformal / copy creation                FLASH_Obj.c:406-636
 *   FLASH_Obj_free{,_without_buffer,_hierarchy}, extract_buffer  FLASH_Obj.c:638-763
 *   FLASH_Obj_attach_buffer{,_hierarchy}     FLASH_Obj.c:781-915
 *   depth / blocksizes / base scalar dims    FLASH_Obj.c:14-140
 *   scalar dims and offsets of views         src/base/flamec/hierarchy/main/FLASH_View.c:600-750
 *   FLASH_Copy_flat_to_hier / hier_to_flat   src/base/flamec/hierarchy/main/FLASH_Copy_other.c:81-138
 *
 * A hierarchical object is a tree of FLA_Obj.  An inner node has elemtype
 * FLA_MATRIX, the scalar datatype of the matrix, its size in blocks (m, n)
 * and in scalars (m_inner, n_inner), and a host array of its children
 * (block-column-major, rs = 1, cs = number of block rows).  A leaf is a flat
 * column-major block (elemtype FLA_SCALAR) whose buffer lies in one device
 * allocation shared by the whole tree: the leaves take consecutive segments
 * in the order the reference assigns them (block-column-major at every level,
 * each leaf's leading dimension equal to its row count), so
 * FLASH_Obj_extract_buffer returns the same layout a libflame user sees.
 * Objects made by FLASH_Obj_attach_buffer instead point their leaves into a
 * column-major matrix (leading dimension = its column stride).
 *
 * Data moves between the leaves and a flat column-major matrix in one
 * tile-copy kernel launch (fla_hip_tile_copy), never leaf by leaf from the
 * host.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "FLAME.h"
#include "FLA_hip.h"
#include "fla_internal.h"

/* ---- queries --------------------------------------------------------------- */

FLA_Datatype FLASH_Obj_datatype(FLA_Obj H) { return FLA_Obj_datatype(H); }

/* FLASH_Obj.c:20-47: levels down the top-left path. */
dim_t FLASH_Obj_depth(FLA_Obj H) {
  dim_t depth = 0;
  FLA_Elemtype et = FLA_Obj_elemtype(H);
  FLA_Obj* buf = (FLA_Obj*)FLA_Obj_base_buffer(H);
  while (et == FLA_MATRIX && buf != NULL) {
    ++depth;
    et = FLA_Obj_elemtype(buf[0]);
    buf = (FLA_Obj*)FLA_Obj_base_buffer(buf[0]);
  }
  return depth;
}

/* FLASH_Obj.c:49-81: block sizes (in scalars, of the top-left child's base)
   per level. */
dim_t FLASH_Obj_blocksizes(FLA_Obj H, dim_t* b_m, dim_t* b_n) {
  dim_t depth = 0;
  FLA_Elemtype et = FLA_Obj_elemtype(H);
  FLA_Obj* buf = (FLA_Obj*)FLA_Obj_base_buffer(H);
  while (et == FLA_MATRIX && buf != NULL) {
    b_m[depth] = FLA_Obj_base_length(buf[0]);
    b_n[depth] = FLA_Obj_base_width(buf[0]);
    ++depth;
    et = FLA_Obj_elemtype(buf[0]);
    buf = (FLA_Obj*)FLA_Obj_base_buffer(buf[0]);
  }
  return depth;
}

/* Scalar rows / columns of a stored child (its m_inner / n_inner). */
static dim_t child_rows(FLA_Obj c) {
  return FLA_Obj_elemtype(c) == FLA_SCALAR ? c.m : c.m_inner;
}
static dim_t child_cols(FLA_Obj c) {
  return FLA_Obj_elemtype(c) == FLA_SCALAR ? c.n : c.n_inner;
}

static FLA_Obj child_at(FLA_Obj H, dim_t i, dim_t j) {
  const FLA_Obj* buf = (const FLA_Obj*)H.base->buffer;
  return buf[(H.offm + i) * H.base->rs + (H.offn + j) * H.base->cs];
}

/* FLASH_Obj.c:83-140: the whole base object, not the view. */
dim_t FLASH_Obj_base_scalar_length(FLA_Obj H) {
  dim_t i, m = 0;
  const FLA_Obj* buf;
  if (FLA_Obj_elemtype(H) == FLA_SCALAR) return FLA_Obj_base_length(H);
  buf = (const FLA_Obj*)H.base->buffer;
  for (i = 0; i < H.base->m; ++i) m += buf[i * H.base->rs].base->m_inner;
  return m;
}

dim_t FLASH_Obj_base_scalar_width(FLA_Obj H) {
  dim_t j, n = 0;
  const FLA_Obj* buf;
  if (FLA_Obj_elemtype(H) == FLA_SCALAR) return FLA_Obj_base_width(H);
  buf = (const FLA_Obj*)H.base->buffer;
  for (j = 0; j < H.base->n; ++j) n += buf[j * H.base->cs].base->n_inner;
  return n;
}

/* FLASH_View.c:600-672: sums over the view's block column 0 / block row 0. */
dim_t FLASH_Obj_scalar_length(FLA_Obj H) {
  dim_t i, m = 0;
  if (FLA_Obj_elemtype(H) == FLA_SCALAR) return H.m;
  if (H.m == 0 || H.n == 0) return 0;
  for (i = 0; i < H.m; ++i) m += child_rows(child_at(H, i, 0));
  return m;
}

dim_t FLASH_Obj_scalar_width(FLA_Obj H) {
  dim_t j, n = 0;
  if (FLA_Obj_elemtype(H) == FLA_SCALAR) return H.n;
  if (H.m == 0 || H.n == 0) return 0;
  for (j = 0; j < H.n; ++j) n += child_cols(child_at(H, 0, j));
  return n;
}

dim_t FLASH_Obj_scalar_min_dim(FLA_Obj H) {
  dim_t m = FLASH_Obj_scalar_length(H), n = FLASH_Obj_scalar_width(H);
  return m < n ? m : n;
}
dim_t FLASH_Obj_scalar_max_dim(FLA_Obj H) {
  dim_t m = FLASH_Obj_scalar_length(H), n = FLASH_Obj_scalar_width(H);
  return m > n ? m : n;
}
dim_t FLASH_Obj_scalar_vector_dim(FLA_Obj H) {
  return FLASH_Obj_scalar_length(H) == 1 ? FLASH_Obj_scalar_width(H) : FLASH_Obj_scalar_length(H);
}

dim_t FLASH_Obj_scalar_length_tl(FLA_Obj H) {
  if (FLA_Obj_elemtype(H) == FLA_SCALAR) return FLA_Obj_base_length(H);
  return FLASH_Obj_base_scalar_length(((FLA_Obj*)FLA_Obj_base_buffer(H))[0]);
}
dim_t FLASH_Obj_scalar_width_tl(FLA_Obj H) {
  if (FLA_Obj_elemtype(H) == FLA_SCALAR) return FLA_Obj_base_width(H);
  return FLASH_Obj_base_scalar_width(((FLA_Obj*)FLA_Obj_base_buffer(H))[0]);
}

/* FLASH_View.c:693-721: offset in blocks times the top-left block size, plus
   the offset inside the first block. */
dim_t FLASH_Obj_scalar_row_offset(FLA_Obj H) {
  if (FLA_Obj_elemtype(H) == FLA_SCALAR) return H.offm;
  return H.offm * FLASH_Obj_scalar_length_tl(H) +
         FLASH_Obj_scalar_row_offset(*FLASH_OBJ_PTR_AT(H));
}
dim_t FLASH_Obj_scalar_col_offset(FLA_Obj H) {
  if (FLA_Obj_elemtype(H) == FLA_SCALAR) return H.offn;
  return H.offn * FLASH_Obj_scalar_width_tl(H) +
         FLASH_Obj_scalar_col_offset(*FLASH_OBJ_PTR_AT(H));
}

/* ---- creation -------------------------------------------------------------- */

static FLA_Error node_create(FLA_Datatype dt, dim_t num_m, dim_t num_n, dim_t m_inner,
                             dim_t n_inner, FLA_Obj* H) {
  FLA_Base_obj* b = (FLA_Base_obj*)calloc(1, sizeof(FLA_Base_obj));
  if (b == NULL) return FLA_FAILURE;
  b->datatype = dt;
  b->elemtype = FLA_MATRIX;
  b->m = num_m;
  b->n = num_n;
  b->rs = 1;
  b->cs = num_m > 0 ? num_m : 1;
  b->m_inner = m_inner;
  b->n_inner = n_inner;
  b->n_elem_alloc = num_m * num_n;
  b->uplo = FLA_FULL_MATRIX;
  b->buffer_info = FLA_BUFFER_HOST;
  if (num_m * num_n > 0) {
    b->buffer = calloc(num_m * num_n, sizeof(FLA_Obj));
    if (b->buffer == NULL) {
      free(b);
      return FLA_FAILURE;
    }
  }
  memset(H, 0, sizeof(*H));
  H->m = num_m;
  H->n = num_n;
  H->m_inner = m_inner;
  H->n_inner = n_inner;
  H->base = b;
  return FLA_SUCCESS;
}

/* Builds the tree for an m x n matrix (FLASH_Obj.c:271-404).  elem_m/elem_n
   hold the scalar extent of a full block at each remaining level; each leaf
   takes the next m*n elements of the flat allocation at *cursor (NULL: no
   buffer).  m_index / n_index locate leaves in units of storage blocks. */
static FLA_Error hier_build(FLA_Datatype dt, dim_t m, dim_t n, dim_t depth, const dim_t* elem_m,
                            const dim_t* elem_n, char** cursor, size_t esz, dim_t mi0, dim_t ni0,
                            const dim_t* dsz_m, const dim_t* dsz_n, FLA_Obj* H) {
  if (depth == 0) {
    FLA_Error e = FLA_Obj_create_without_buffer(dt, m, n, H);
    if (e != FLA_SUCCESS) return e;
    H->base->m_index = mi0;
    H->base->n_index = ni0;
    if (*cursor != NULL) {
      H->base->buffer = *cursor;
      H->base->rs = 1;
      H->base->cs = m > 0 ? m : 1;
      H->base->buffer_info = FLA_BUFFER_DEVICE | FLA_BUFFER_ATTACHED | FLA_BUFFER_TREE;
      *cursor += (size_t)m * (size_t)n * esz;
    }
    return FLA_SUCCESS;
  } else {
    const dim_t bm = elem_m[0], bn = elem_n[0];
    const dim_t num_m = (m + bm - 1) / bm, num_n = (n + bn - 1) / bn;
    dim_t i, j;
    FLA_Obj* buf;
    FLA_Error e = node_create(dt, num_m, num_n, m, n, H);
    if (e != FLA_SUCCESS) return e;
    buf = (FLA_Obj*)H->base->buffer;
    for (j = 0; j < num_n; ++j) {
      const dim_t nn = (j != num_n - 1 || n % bn == 0) ? bn : n % bn;
      for (i = 0; i < num_m; ++i) {
        const dim_t mm = (i != num_m - 1 || m % bm == 0) ? bm : m % bm;
        e = hier_build(dt, mm, nn, depth - 1, elem_m + 1, elem_n + 1, cursor, esz,
                       mi0 + i * dsz_m[0], ni0 + j * dsz_n[0], dsz_m + 1, dsz_n + 1,
                       &buf[j * num_m + i]);
        if (e != FLA_SUCCESS) return e;
      }
    }
    return FLA_SUCCESS;
  }
}

static FLA_Error flash_create(int with_buffer, FLA_Datatype datatype, dim_t m, dim_t n,
                              dim_t depth, const dim_t* b_m, const dim_t* b_n, FLA_Obj* H) {
  dim_t elem_m[16], elem_n[16], dsz_m[16], dsz_n[16], i;
  char* cursor = NULL;
  size_t esz;
  FLA_Error e;
  if (depth == 0) {
    return with_buffer ? FLA_Obj_create(datatype, m, n, 0, 0, H)
                       : FLA_Obj_create_without_buffer(datatype, m, n, H);
  }
  if (depth > 16 || b_m == NULL || b_n == NULL) return FLA_FAILURE;
  for (i = 0; i < depth; ++i)
    if (b_m[i] == 0 || b_n[i] == 0) {
      fprintf(stderr, "libflame_amd: FLASH_Obj_create: zero blocksize at level %lu\n",
              (unsigned long)i);
      return FLA_FAILURE;
    }
  /* elem_m[i]: scalar rows of a full block at level i; dsz_m[i]: storage
     blocks (innermost size) per block at level i (FLASH_Obj.c:200-214). */
  elem_m[depth - 1] = b_m[depth - 1];
  elem_n[depth - 1] = b_n[depth - 1];
  dsz_m[depth - 1] = dsz_n[depth - 1] = 1;
  for (i = depth - 1; i > 0; --i) {
    elem_m[i - 1] = elem_m[i] * b_m[i - 1];
    elem_n[i - 1] = elem_n[i] * b_n[i - 1];
    dsz_m[i - 1] = dsz_m[i] * b_m[i - 1];
    dsz_n[i - 1] = dsz_n[i] * b_n[i - 1];
  }
  {
    FLA_Obj probe;
    if (FLA_Obj_create_without_buffer(datatype, 1, 1, &probe) != FLA_SUCCESS) return FLA_FAILURE;
    esz = (size_t)FLA_Obj_elem_size(probe);
    FLA_Obj_free_without_buffer(&probe);
  }
  if (with_buffer && m * n > 0) {
    void* p = NULL;
    int st = fla_hip_malloc(&p, (size_t)m * (size_t)n * esz);
    if (st != FLA_HIP_OK || p == NULL) {
      fprintf(stderr, "libflame_amd: FLASH_Obj_create: device allocation failed (%s)\n",
              fla_hip_error_string(st));
      return FLA_GPU_ERROR;
    }
    cursor = (char*)p;
  }
  e = hier_build(datatype, m, n, depth, elem_m, elem_n, &cursor, esz, 0, 0, dsz_m, dsz_n, H);
  return e;
}

FLA_Error FLASH_Obj_create(FLA_Datatype datatype, dim_t m, dim_t n, dim_t depth, dim_t* b_mn,
                           FLA_Obj* H) {
  return flash_create(1, datatype, m, n, depth, b_mn, b_mn, H);
}
FLA_Error FLASH_Obj_create_ext(FLA_Datatype datatype, dim_t m, dim_t n, dim_t depth, dim_t* b_m,
                               dim_t* b_n, FLA_Obj* H) {
  return flash_create(1, datatype, m, n, depth, b_m, b_n, H);
}
FLA_Error FLASH_Obj_create_without_buffer(FLA_Datatype datatype, dim_t m, dim_t n, dim_t depth,
                                          dim_t* b_mn, FLA_Obj* H) {
  return flash_create(0, datatype, m, n, depth, b_mn, b_mn, H);
}
FLA_Error FLASH_Obj_create_without_buffer_ext(FLA_Datatype datatype, dim_t m, dim_t n,
                                              dim_t depth, dim_t* b_m, dim_t* b_n, FLA_Obj* H) {
  return flash_create(0, datatype, m, n, depth, b_m, b_n, H);
}

/* FLASH_Obj.c:406-457.  Only the whole object is supported as the model (a
   view would need the reference's view adjustment of the new hierarchy). */
FLA_Error FLASH_Obj_create_conf_to(FLA_Trans trans, FLA_Obj H, FLA_Obj* H_new) {
  dim_t b_m[16], b_n[16], depth;
  if (FLA_Obj_elemtype(H) == FLA_SCALAR)
    return FLA_Obj_create(FLA_Obj_datatype(H), trans == FLA_NO_TRANSPOSE ? H.m : H.n,
                          trans == FLA_NO_TRANSPOSE ? H.n : H.m, 0, 0, H_new);
  if (trans == FLA_TRANSPOSE || trans == FLA_CONJ_TRANSPOSE) {
    fprintf(stderr, "libflame_amd: FLASH_Obj_create_conf_to: transposition not implemented\n");
    return FLA_FAILURE;
  }
  if (H.offm != 0 || H.offn != 0 || H.m != H.base->m || H.n != H.base->n) {
    fprintf(stderr, "libflame_amd: FLASH_Obj_create_conf_to: H must be a whole object\n");
    return FLA_FAILURE;
  }
  depth = FLASH_Obj_blocksizes(H, b_m, b_n);
  return FLASH_Obj_create_ext(FLA_Obj_datatype(H), FLASH_Obj_base_scalar_length(H),
                              FLASH_Obj_base_scalar_width(H), depth, b_m, b_n, H_new);
}

FLA_Error FLASH_Obj_create_hier_conf_to_flat(FLA_Trans trans, FLA_Obj F, dim_t depth,
                                             dim_t* b_mn, FLA_Obj* H) {
  const int t = (trans == FLA_TRANSPOSE || trans == FLA_CONJ_TRANSPOSE);
  return FLASH_Obj_create(FLA_Obj_datatype(F), t ? F.n : F.m, t ? F.m : F.n, depth, b_mn, H);
}
FLA_Error FLASH_Obj_create_hier_conf_to_flat_ext(FLA_Trans trans, FLA_Obj F, dim_t depth,
                                                 dim_t* b_m, dim_t* b_n, FLA_Obj* H) {
  const int t = (trans == FLA_TRANSPOSE || trans == FLA_CONJ_TRANSPOSE);
  return FLASH_Obj_create_ext(FLA_Obj_datatype(F), t ? F.n : F.m, t ? F.m : F.n, depth, b_m,
                              b_n, H);
}
FLA_Error FLASH_Obj_create_flat_conf_to_hier(FLA_Trans trans, FLA_Obj H, FLA_Obj* F) {
  const int t = (trans == FLA_TRANSPOSE || trans == FLA_CONJ_TRANSPOSE);
  const dim_t m = FLASH_Obj_scalar_length(H), n = FLASH_Obj_scalar_width(H);
  return FLA_Obj_create(FLASH_Obj_datatype(H), t ? n : m, t ? m : n, 0, 0, F);
}

FLA_Error FLASH_Obj_create_copy_of(FLA_Trans trans, FLA_Obj H_cur, FLA_Obj* H_new) {
  FLA_Error e = FLASH_Obj_create_conf_to(trans, H_cur, H_new);
  if (e != FLA_SUCCESS) return e;
  return FLASH_Copy(H_cur, *H_new);
}

FLA_Error FLASH_Obj_create_hier_copy_of_flat(FLA_Obj F, dim_t depth, dim_t* b_mn, FLA_Obj* H) {
  FLA_Error e = FLASH_Obj_create_hier_conf_to_flat(FLA_NO_TRANSPOSE, F, depth, b_mn, H);
  if (e != FLA_SUCCESS) return e;
  return FLASH_Copy_flat_to_hier(F, 0, 0, *H);
}
FLA_Error FLASH_Obj_create_hier_copy_of_flat_ext(FLA_Obj F, dim_t depth, dim_t* b_m,
                                                 dim_t* b_n, FLA_Obj* H) {
  FLA_Error e = FLASH_Obj_create_hier_conf_to_flat_ext(FLA_NO_TRANSPOSE, F, depth, b_m, b_n, H);
  if (e != FLA_SUCCESS) return e;
  return FLASH_Copy_flat_to_hier(F, 0, 0, *H);
}
FLA_Error FLASH_Obj_create_flat_copy_of_hier(FLA_Obj H, FLA_Obj* F) {
  FLA_Error e = FLASH_Obj_create_flat_conf_to_hier(FLA_NO_TRANSPOSE, H, F);
  if (e != FLA_SUCCESS) return e;
  return FLASH_Copy_hier_to_flat(0, 0, H, *F);
}

/* ---- release --------------------------------------------------------------- */

/* FLASH_Obj.c:741-763: the buffer of the top-left leaf (the start of the
   shared allocation for objects made by FLASH_Obj_create). */
void* FLASH_Obj_extract_buffer(FLA_Obj H) {
  FLA_Elemtype et = FLA_Obj_elemtype(H);
  FLA_Obj* buf = (FLA_Obj*)FLA_Obj_base_buffer(H);
  while (et == FLA_MATRIX && buf != NULL) {
    et = FLA_Obj_elemtype(buf[0]);
    buf = (FLA_Obj*)FLA_Obj_base_buffer(buf[0]);
  }
  return (void*)buf;
}

void FLASH_Obj_free_hierarchy(FLA_Obj* H) {
  if (H->base == NULL) return;
  if (FLA_Obj_elemtype(*H) == FLA_SCALAR) {
    FLA_Obj_free_without_buffer(H);
    return;
  } else {
    FLA_Obj* buf = (FLA_Obj*)H->base->buffer;
    dim_t i;
    for (i = 0; i < H->base->n_elem_alloc; ++i) FLASH_Obj_free_hierarchy(&buf[i]);
    free(buf);
    free(H->base);
    memset(H, 0, sizeof(*H));
  }
}

void FLASH_Obj_free(FLA_Obj* H) {
  if (FLA_Obj_elemtype(*H) == FLA_MATRIX) {
    /* The leaves of a tree made by FLASH_Obj_create share one allocation that
       starts at the top-left leaf and is owned by the tree; leaves attached
       to a caller's matrix own nothing. */
    FLA_Obj* buf = (FLA_Obj*)FLA_Obj_base_buffer(*H);
    while (buf != NULL && FLA_Obj_elemtype(buf[0]) == FLA_MATRIX)
      buf = (FLA_Obj*)FLA_Obj_base_buffer(buf[0]);
    if (buf != NULL && buf[0].base != NULL && (buf[0].base->buffer_info & FLA_BUFFER_TREE) &&
        buf[0].base->buffer != NULL)
      fla_hip_free(buf[0].base->buffer);
    FLASH_Obj_free_hierarchy(H);
  } else {
    FLA_Obj_free(H);
  }
}

void FLASH_Obj_free_without_buffer(FLA_Obj* H) {
  if (FLA_Obj_elemtype(*H) == FLA_MATRIX)
    FLASH_Obj_free_hierarchy(H);
  else
    FLA_Obj_free_without_buffer(H);
}

/* ---- leaves, flat copies, attached buffers ----------------------------------- */

typedef struct {
  long long* d; /* 6 per leaf: pointer, ld, rows, cols, row / column in the region */
  size_t count, cap;
  FLA_Error err;
} leafvec;

static void leafvec_push(leafvec* v, const char* ptr, dim_t ld, dim_t rows, dim_t cols, dim_t fr,
                         dim_t fc) {
  if (v->count == v->cap) {
    size_t cap = v->cap ? 2 * v->cap : 64;
    long long* d = (long long*)realloc(v->d, cap * 6 * sizeof(long long));
    if (d == NULL) {
      v->err = FLA_FAILURE;
      return;
    }
    v->d = d;
    v->cap = cap;
  }
  {
    long long* e = v->d + 6 * v->count++;
    e[0] = (long long)(size_t)ptr;
    e[1] = (long long)ld;
    e[2] = (long long)rows;
    e[3] = (long long)cols;
    e[4] = (long long)fr;
    e[5] = (long long)fc;
  }
}

/* Leaves of H (whose scalar (0,0) sits at (r0, c0)) clipped to the region
   [i0, i0 + m) x [j0, j0 + n); offsets of each piece relative to (i0, j0). */
static void collect(FLA_Obj H, dim_t r0, dim_t c0, dim_t i0, dim_t j0, dim_t m, dim_t n,
                    dim_t esz, leafvec* v) {
  if (v->err != FLA_SUCCESS || H.m == 0 || H.n == 0) return;
  if (FLA_Obj_elemtype(H) == FLA_SCALAR) {
    const dim_t ra = r0 > i0 ? r0 : i0, rb = (r0 + H.m < i0 + m) ? r0 + H.m : i0 + m;
    const dim_t ca = c0 > j0 ? c0 : j0, cb = (c0 + H.n < j0 + n) ? c0 + H.n : j0 + n;
    if (ra >= rb || ca >= cb) return;
    if (H.base->buffer == NULL || !(H.base->buffer_info & FLA_BUFFER_DEVICE) ||
        (H.base->rs != 1 && H.m > 1)) {
      fprintf(stderr,
              "libflame_amd: FLASH leaves must be column-major blocks in device memory\n");
      v->err = FLA_FAILURE;
      return;
    }
    leafvec_push(v,
                 (const char*)FLA_Obj_buffer_at_view(H) +
                     ((ra - r0) + (ca - c0) * H.base->cs) * esz,
                 H.base->cs, rb - ra, cb - ca, ra - i0, ca - j0);
    return;
  } else {
    dim_t i, j, cacc = c0;
    for (j = 0; j < H.n && cacc < j0 + n; ++j) {
      const dim_t w = child_cols(child_at(H, 0, j));
      if (cacc + w > j0) {
        dim_t racc = r0;
        for (i = 0; i < H.m && racc < i0 + m; ++i) {
          const FLA_Obj c = child_at(H, i, j);
          const dim_t h = child_rows(c);
          if (racc + h > i0) collect(c, racc, cacc, i0, j0, m, n, esz, v);
          racc += h;
        }
      }
      cacc += w;
    }
  }
}

/* Scalar element size of a (possibly hierarchical) object. */
static dim_t scalar_esz(FLA_Obj H) { return fla_datatype_size(FLA_Obj_datatype(H)); }

/* Copies between the leaves covering the scalar region (i, j, m, n) of the
   hierarchical view H and the device matrix (dev, ld): to_leaves != 0 writes
   the leaves, else reads them.  One tile-copy launch. */
FLA_Error fla_hier_copy(FLA_Obj H, dim_t i, dim_t j, dim_t m, dim_t n, void* dev, int ld,
                        int to_leaves) {
  leafvec v = {NULL, 0, 0, FLA_SUCCESS};
  int st;
  if (m == 0 || n == 0) return FLA_SUCCESS;
  collect(H, 0, 0, i, j, m, n, scalar_esz(H), &v);
  if (v.err != FLA_SUCCESS) {
    free(v.d);
    return v.err;
  }
  st = fla_hip_tile_copy((int)v.count, v.d, dev, (long)ld, (int)scalar_esz(H), to_leaves);
  free(v.d);
  if (st != FLA_HIP_OK) {
    fprintf(stderr, "libflame_amd: FLASH tile copy failed: %s\n", fla_hip_error_string(st));
    return FLA_GPU_ERROR;
  }
  return FLA_SUCCESS;
}

/* FLASH_Copy_other.c:81-138: F is copied to / from H's scalar submatrix at
   (i, j) of F's size. */
FLA_Error FLASH_Copy_flat_to_hier(FLA_Obj F, dim_t i, dim_t j, FLA_Obj H) {
  fla_devmat dF;
  FLA_Error e;
  if (FLA_Obj_elemtype(H) == FLA_SCALAR) {
    FLA_Obj Hs, t0, t1, t2, t3;
    FLA_Part_2x2(H, &t0, &t1, &t2, &Hs, i, j, FLA_TL);
    FLA_Part_2x2(Hs, &t0, &t1, &t2, &t3, F.m, F.n, FLA_TL);
    return FLA_Copy(F, t0);
  }
  if ((e = fla_devmat_acquire(F, 1, &dF)) != FLA_SUCCESS) return e;
  e = fla_hier_copy(H, i, j, F.m, F.n, dF.dev, dF.ld, 1);
  if (e == FLA_SUCCESS && fla_hip_synchronize() != FLA_HIP_OK) e = FLA_GPU_ERROR;
  fla_devmat_release(&dF, 0);
  return e;
}

FLA_Error FLASH_Copy_hier_to_flat(dim_t i, dim_t j, FLA_Obj H, FLA_Obj F) {
  fla_devmat dF;
  FLA_Error e, e2;
  if (FLA_Obj_elemtype(H) == FLA_SCALAR) {
    FLA_Obj Hs, t0, t1, t2, t3;
    FLA_Part_2x2(H, &t0, &t1, &t2, &Hs, i, j, FLA_TL);
    FLA_Part_2x2(Hs, &t0, &t1, &t2, &t3, F.m, F.n, FLA_TL);
    return FLA_Copy(t0, F);
  }
  if ((e = fla_devmat_acquire(F, 0, &dF)) != FLA_SUCCESS) return e;
  e = fla_hier_copy(H, i, j, F.m, F.n, dF.dev, dF.ld, 0);
  if (e == FLA_SUCCESS && fla_hip_synchronize() != FLA_HIP_OK) e = FLA_GPU_ERROR;
  e2 = fla_devmat_release(&dF, e == FLA_SUCCESS);
  return e != FLA_SUCCESS ? e : e2;
}

FLA_Error FLASH_Obj_flatten(FLA_Obj H, FLA_Obj F) { return FLASH_Copy_hier_to_flat(0, 0, H, F); }
FLA_Error FLASH_Obj_hierarchify(FLA_Obj F, FLA_Obj H) {
  return FLASH_Copy_flat_to_hier(F, 0, 0, H);
}

/* B := A for conformal objects, either of them hierarchical (FLA_Copy moves
   hierarchical views through flat device copies of their scalar extent). */
FLA_Error FLASH_Copy(FLA_Obj A, FLA_Obj B) { return FLA_Copy(A, B); }

/* Points every leaf of H (scalar (0,0) at (r0, c0)) into the matrix at
   `base` with strides rs, cs (FLASH_Obj.c:815-915). */
static void attach_walk(FLA_Obj H, dim_t r0, dim_t c0, char* base, dim_t rs, dim_t cs,
                        dim_t esz, int where) {
  if (FLA_Obj_elemtype(H) == FLA_SCALAR) {
    H.base->buffer = base + (r0 * rs + c0 * cs) * esz;
    H.base->rs = rs;
    H.base->cs = cs;
    H.base->buffer_info = where | FLA_BUFFER_ATTACHED;
    return;
  } else {
    dim_t i, j, cacc = c0;
    for (j = 0; j < H.base->n; ++j) {
      const FLA_Obj* buf = (const FLA_Obj*)H.base->buffer;
      dim_t racc = r0;
      for (i = 0; i < H.base->m; ++i) {
        const FLA_Obj c = buf[i * H.base->rs + j * H.base->cs];
        attach_walk(c, racc, cacc, base, rs, cs, esz, where);
        racc += child_rows(c);
      }
      cacc += child_cols(buf[j * H.base->cs]);
    }
  }
}

FLA_Error FLASH_Obj_attach_buffer(void* buffer, dim_t rs, dim_t cs, FLA_Obj* H) {
  if (FLA_Obj_elemtype(*H) == FLA_SCALAR) return FLA_Obj_attach_buffer(buffer, rs, cs, H);
  if (rs == 0 && cs == 0) {
    rs = 1;
    cs = FLASH_Obj_base_scalar_length(*H);
  }
  attach_walk(*H, 0, 0, (char*)buffer, rs, cs, scalar_esz(*H),
              fla_hip_is_device_ptr(buffer) ? FLA_BUFFER_DEVICE : FLA_BUFFER_HOST);
  return FLA_SUCCESS;
}

FLA_Error FLASH_Obj_attach_buffer_hierarchy(FLA_Obj F, FLA_Obj* H) {
  return FLASH_Obj_attach_buffer(FLA_Obj_buffer_at_view(F), F.base->rs, F.base->cs, H);
}
