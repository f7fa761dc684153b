// KnnLeaf2Leaf on float positions in dims 4-10 (the kernel is in knn_nd.h).
#include "knn_nd.h"

namespace jzt {

int knn_leaf_nd_f32(JZT_KNN_ND_ARGS) {
#define JZT_ND_DIM(D)                                                                                           \
  case D:                                                                                                       \
    return knn_leaf_nd_impl<D, float>(s, il_spl, il_src, il_r2, splT, xT, splQ, xQ, src_ids, dst_rows, init_r2, retry, \
                                 rad, ids, tmp, tmp_bytes, nleaves_q, npart_q, out_rows, il_size, k, boxsize);
  switch (dim) {
    JZT_ND_DIM(4)
    JZT_ND_DIM(5)
    JZT_ND_DIM(6)
    JZT_ND_DIM(7)
    JZT_ND_DIM(8)
    JZT_ND_DIM(9)
    JZT_ND_DIM(10)
    default:
      break;
  }
#undef JZT_ND_DIM
  return fail("KnnLeaf2Leaf: unsupported dim " + std::to_string(dim) + " for float positions");
}

}  // namespace jzt
