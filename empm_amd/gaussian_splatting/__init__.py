"""Drop-in mirror of the reference's third_party/gaussian_splatting motion
interpolation (dynamic_utils.py) on the HIP kernels of libempm."""
from .dynamic_utils import (LBSInterpolator, calc_weights_vals_from_indices, create_relation_matrix,
                            get_topk_indices, interpolate_motions_speedup, knn_weights, knn_weights_sparse, mat2quat,
                            quat2mat)

__all__ = ["LBSInterpolator", "calc_weights_vals_from_indices", "create_relation_matrix", "get_topk_indices",
           "interpolate_motions_speedup", "knn_weights", "knn_weights_sparse", "mat2quat", "quat2mat"]
