"""ctypes view of the C ABI declared in include/whatprot_amd.h.

Only layouts and prototypes live here; the host-side mirror of the reference's
classes is in whatprot_amd/__init__.py.
"""
import ctypes
import os

WP_MAX_CHANNELS = 8
WP_MAX_TIMESTEPS = 64
WP_MAX_DYES = 15

WP_OK = 0
WP_PRECISION_FP64 = 0
WP_PRECISION_FP32 = 1

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libwhatprot_amd.so")


class ChannelModelC(ctypes.Structure):
    _fields_ = [
        ("p_bleach", ctypes.c_double),
        ("p_dud", ctypes.c_double),
        ("bg_sig", ctypes.c_double),
        ("mu", ctypes.c_double),
        ("sig", ctypes.c_double),
        ("interactions", ctypes.c_double * WP_MAX_CHANNELS),
        ("flat_interactions", ctypes.c_double * WP_MAX_CHANNELS),
    ]


class ModelC(ctypes.Structure):
    _fields_ = [
        ("num_channels", ctypes.c_int),
        ("p_edman_failure", ctypes.c_double),
        ("p_detach", ctypes.c_double),
        ("p_initial_detach", ctypes.c_double),
        ("p_initial_detach_decay", ctypes.c_double),
        ("p_initial_block", ctypes.c_double),
        ("p_cyclic_block", ctypes.c_double),
        ("channels", ChannelModelC * WP_MAX_CHANNELS),
    ]


class FitSettingsC(ctypes.Structure):
    _fields_ = [
        ("hold_p_edman_failure", ctypes.c_int),
        ("hold_p_detach", ctypes.c_int),
        ("hold_p_initial_detach", ctypes.c_int),
        ("hold_p_initial_detach_decay", ctypes.c_int),
        ("hold_p_initial_block", ctypes.c_int),
        ("hold_p_cyclic_block", ctypes.c_int),
        ("hold_p_bleach", ctypes.c_int * WP_MAX_CHANNELS),
        ("hold_p_dud", ctypes.c_int * WP_MAX_CHANNELS),
    ]


def fit_stats_len(T, C):
    """WP_FIT_STATS_LEN."""
    return 6 + 2 * (T - 1) + 4 * C


_lib = None


class RadiometriesC(ctypes.Structure):
    _fields_ = [("num_timesteps", ctypes.c_int), ("num_channels", ctypes.c_int), ("num_reads", ctypes.c_int),
                ("values", ctypes.POINTER(ctypes.c_double))]


class DyeTracksC(ctypes.Structure):
    _fields_ = [("num_timesteps", ctypes.c_int), ("num_channels", ctypes.c_int), ("num_dye_tracks", ctypes.c_int),
                ("num_sources", ctypes.c_int), ("counts", ctypes.POINTER(ctypes.c_short)),
                ("source_offsets", ctypes.POINTER(ctypes.c_int)), ("source_ids", ctypes.POINTER(ctypes.c_int)),
                ("source_counts", ctypes.POINTER(ctypes.c_int)), ("source_hits", ctypes.POINTER(ctypes.c_int))]


class WhatprotError(RuntimeError):
    pass


def lib():
    """Loads libwhatprot_amd.so. Fails loudly when it is missing: there is no
    CPU fallback for the scoring path."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise WhatprotError(
            "libwhatprot_amd.so is not built (run python whatprot_amd/build.py or "
            "__graft_entry__.build())")
    L = ctypes.CDLL(LIB_PATH)
    c_int, c_uint, c_double, c_void_p = ctypes.c_int, ctypes.c_uint, ctypes.c_double, ctypes.c_void_p
    P = ctypes.POINTER
    L.wp_last_error.restype = ctypes.c_char_p
    L.wp_last_error.argtypes = []
    L.wp_model_with_mu_as_one.restype = None
    L.wp_model_with_mu_as_one.argtypes = [P(ModelC), P(ModelC)]
    L.wp_device_count.restype = c_int
    L.wp_device_count.argtypes = []
    L.wp_hmm_classifier_create.restype = c_int
    L.wp_hmm_classifier_create.argtypes = [c_uint, c_uint, P(ModelC), c_double, c_int,
                                           P(ctypes.c_char_p), P(c_int), P(c_int), c_int, P(c_void_p)]
    L.wp_hmm_classifier_destroy.restype = None
    L.wp_hmm_classifier_destroy.argtypes = [c_void_p]
    L.wp_hmm_classifier_max_num_dyes.restype = c_int
    L.wp_hmm_classifier_max_num_dyes.argtypes = [c_void_p]
    L.wp_hmm_classify.restype = c_int
    L.wp_hmm_classify.argtypes = [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]
    L.wp_hmm_classify_candidates.restype = c_int
    L.wp_hmm_classify_candidates.argtypes = [c_void_p, c_int, c_void_p, c_void_p, c_void_p,
                                             c_void_p, c_void_p, c_void_p, c_void_p]
    L.wp_hmm_classify_device.restype = c_int
    L.wp_hmm_classify_device.argtypes = [c_void_p, c_int, c_void_p, c_void_p, c_void_p,
                                         c_void_p, c_void_p, c_void_p]
    L.wp_hmm_classify_device_noalloc.restype = c_int
    L.wp_hmm_classify_device_noalloc.argtypes = [c_void_p, c_int, c_void_p, c_void_p, c_void_p,
                                                 c_void_p, c_void_p, c_void_p]
    L.wp_hmm_pair_log_probability.restype = c_int
    L.wp_hmm_pair_log_probability.argtypes = [c_void_p, c_int, c_void_p, c_void_p]
    L.wp_hmm_set_timing.restype = c_int
    L.wp_hmm_set_timing.argtypes = [c_void_p, c_int]
    L.wp_hmm_last_timing.restype = c_int
    L.wp_hmm_last_timing.argtypes = [c_void_p, P(c_double), P(c_double)]
    L.wp_hmm_set_variant.restype = c_int
    L.wp_hmm_set_variant.argtypes = [c_void_p, c_int]
    L.wp_hmm_get_variants.restype = c_int
    L.wp_hmm_get_variants.argtypes = [c_void_p, c_void_p, c_void_p, c_int]
    L.wp_hmm_set_graph.restype = c_int
    L.wp_hmm_set_graph.argtypes = [c_void_p, c_int]
    L.wp_hmm_set_branches.restype = c_int
    L.wp_hmm_set_branches.argtypes = [c_void_p, c_int]
    L.wp_read_radiometries.restype = c_int
    L.wp_read_radiometries.argtypes = [ctypes.c_char_p, c_void_p, P(RadiometriesC)]
    L.wp_free_radiometries.restype = None
    L.wp_free_radiometries.argtypes = [P(RadiometriesC)]
    L.wp_write_radiometries.restype = c_int
    L.wp_write_radiometries.argtypes = [ctypes.c_char_p, c_void_p, c_int, c_int, c_int, c_void_p]
    L.wp_read_dye_tracks.restype = c_int
    L.wp_read_dye_tracks.argtypes = [ctypes.c_char_p, P(DyeTracksC)]
    L.wp_free_dye_tracks.restype = None
    L.wp_free_dye_tracks.argtypes = [P(DyeTracksC)]
    L.wp_write_dye_tracks.restype = c_int
    L.wp_write_dye_tracks.argtypes = [ctypes.c_char_p, c_int, c_int, c_int] + [c_void_p] * 5
    L.wp_write_scored_classifications.restype = c_int
    L.wp_write_scored_classifications.argtypes = [ctypes.c_char_p, c_int, c_void_p, c_void_p]
    L.wp_hmm_fitter_set_batched.restype = c_int
    L.wp_hmm_fitter_set_batched.argtypes = [c_void_p, c_int]
    L.wp_hmm_set_refine.restype = c_int
    L.wp_hmm_set_refine.argtypes = [c_void_p, c_int, c_double, c_double]
    L.wp_hmm_refine_stats.restype = c_int
    L.wp_hmm_refine_stats.argtypes = [c_void_p, c_int, c_void_p]
    L.wp_hmm_work.restype = c_int
    L.wp_hmm_work.argtypes = [c_void_p, c_int, c_void_p, c_void_p]
    L.wp_generate_dye_seqs.restype = c_int
    L.wp_generate_dye_seqs.argtypes = [c_int, c_int, c_int, c_int, c_int, c_int, ctypes.c_uint64,
                                       c_void_p, c_int]
    L.wp_simulate_radiometries.restype = c_int
    L.wp_simulate_radiometries.argtypes = [P(ModelC), c_uint, c_uint, c_int, P(ctypes.c_char_p),
                                           c_void_p, c_int, ctypes.c_uint64, c_void_p, c_void_p]
    L.wp_hmm_fitter_create.restype = c_int
    L.wp_hmm_fitter_create.argtypes = [c_uint, c_uint, ctypes.c_char_p, c_int, c_void_p, c_int, P(c_void_p)]
    L.wp_hmm_fitter_destroy.restype = None
    L.wp_hmm_fitter_destroy.argtypes = [c_void_p]
    L.wp_hmm_fitter_estep.restype = c_int
    L.wp_hmm_fitter_estep.argtypes = [c_void_p, P(ModelC), c_void_p, c_void_p]
    L.wp_fit_mstep.restype = c_int
    L.wp_fit_mstep.argtypes = [c_uint, c_uint, ctypes.c_char_p, P(ModelC), c_int, c_void_p,
                               P(FitSettingsC), P(ModelC), P(c_double)]
    L.wp_hmm_fit_bootstrap.restype = c_int
    L.wp_hmm_fit_bootstrap.argtypes = [c_void_p, P(ModelC), P(FitSettingsC), c_double, c_double, c_int, c_double,
                                       ctypes.c_uint64, c_void_p, c_void_p, P(ModelC), P(c_double), P(c_double),
                                       P(ModelC), P(ModelC), P(c_double)]
    L.wp_fit_bootstrap_counts.restype = c_int
    L.wp_fit_bootstrap_counts.argtypes = [c_int, ctypes.c_uint64, c_int, c_void_p]
    L.wp_hmm_fit.restype = c_int
    L.wp_hmm_fit.argtypes = [c_void_p, P(ModelC), P(FitSettingsC), c_double, c_double, c_int, P(ModelC),
                             P(c_double), P(c_double), P(c_int)]
    L.wp_kdtree_order.restype = c_int
    L.wp_kdtree_order.argtypes = [c_int, c_int, c_int, c_void_p, c_void_p, P(c_int)]
    L.wp_nn_classifier_create.restype = c_int
    L.wp_nn_classifier_create.argtypes = [c_uint, c_uint, P(ModelC), c_int, c_double, c_int, c_void_p, c_void_p,
                                          c_void_p, c_void_p, c_void_p, P(c_void_p)]
    L.wp_nn_classifier_destroy.restype = None
    L.wp_nn_classifier_destroy.argtypes = [c_void_p]
    L.wp_nn_search.restype = c_int
    L.wp_nn_search.argtypes = [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p]
    L.wp_nn_classify.restype = c_int
    L.wp_nn_classify.argtypes = [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p]
    L.wp_nn_classify_top.restype = c_int
    L.wp_nn_classify_top.argtypes = [c_void_p, c_int, c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p]
    L.wp_hybrid_classifier_create.restype = c_int
    L.wp_hybrid_classifier_create.argtypes = [c_uint, c_uint, P(ModelC), c_double, c_int, c_double, c_int,
                                              c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int,
                                              P(ctypes.c_char_p), c_void_p, c_void_p, c_int, P(c_void_p)]
    L.wp_hybrid_classifier_destroy.restype = None
    L.wp_hybrid_classifier_destroy.argtypes = [c_void_p]
    L.wp_hybrid_classify.restype = c_int
    L.wp_hybrid_classify.argtypes = [c_void_p, c_int, c_void_p, c_void_p, c_void_p]
    L.wp_simulate_dye_tracks.restype = c_int
    L.wp_simulate_dye_tracks.argtypes = [P(ModelC), c_uint, c_uint, c_int, P(ctypes.c_char_p), c_void_p, c_void_p,
                                         c_int, ctypes.c_uint64, c_int, c_int, P(c_int), P(c_int), c_void_p, c_void_p,
                                         c_void_p, c_void_p, c_void_p]
    L.wp_simulate_radiometries_gpu.restype = c_int
    L.wp_simulate_radiometries_gpu.argtypes = [P(ModelC), c_uint, c_uint, c_int, P(ctypes.c_char_p),
                                           c_void_p, c_int, ctypes.c_uint64, c_void_p, c_void_p]
    L.wp_simulate_dye_tracks_gpu.restype = c_int
    L.wp_simulate_dye_tracks_gpu.argtypes = [P(ModelC), c_uint, c_uint, c_int, P(ctypes.c_char_p), c_void_p, c_void_p,
                                         c_int, ctypes.c_uint64, c_int, c_int, P(c_int), P(c_int), c_void_p, c_void_p,
                                         c_void_p, c_void_p, c_void_p]
    L.wp_hmm_classifier_num_dye_seqs.restype = c_int
    L.wp_hmm_classifier_num_dye_seqs.argtypes = [c_void_p]
    L.wp_mlp_create.restype = c_int
    L.wp_mlp_create.argtypes = [c_int, c_int, c_int, c_int, ctypes.c_uint64, P(c_void_p)]
    L.wp_mlp_destroy.restype = None
    L.wp_mlp_destroy.argtypes = [c_void_p]
    L.wp_mlp_layer_shape.restype = c_int
    L.wp_mlp_layer_shape.argtypes = [c_void_p, c_int, P(c_int), P(c_int)]
    L.wp_mlp_get_layer.restype = c_int
    L.wp_mlp_get_layer.argtypes = [c_void_p, c_int, c_void_p, c_void_p]
    L.wp_mlp_set_layer.restype = c_int
    L.wp_mlp_set_layer.argtypes = [c_void_p, c_int, c_void_p, c_void_p]
    L.wp_mlp_logits.restype = c_int
    L.wp_mlp_logits.argtypes = [c_void_p, c_int, c_void_p, c_void_p]
    L.wp_mlp_gradients.restype = c_int
    L.wp_mlp_gradients.argtypes = [c_void_p, c_int, c_void_p, c_void_p, P(c_double), c_void_p, c_void_p]
    L.wp_mlp_train.restype = c_int
    L.wp_mlp_train.argtypes = [c_void_p, c_int, c_void_p, c_void_p, c_int, c_int, ctypes.c_float, ctypes.c_uint64,
                               c_void_p]
    L.wp_mlp_shortlist.restype = c_int
    L.wp_mlp_shortlist.argtypes = [c_void_p, c_int, c_void_p, c_int, c_void_p, c_void_p, c_void_p]
    L.wp_shortlist_classify.restype = c_int
    L.wp_shortlist_classify.argtypes = [c_void_p, c_void_p, c_int, c_void_p, c_int, c_double, c_int, c_void_p,
                                        c_void_p, P(c_int)]
    L.wp_shortlist_classify_device.restype = c_int
    L.wp_shortlist_classify_device.argtypes = [c_void_p, c_void_p, c_int, c_void_p, c_int, c_double, c_int,
                                               c_void_p, c_void_p, P(c_int), c_void_p]
    L.wp_fit_offset_exponential.restype = c_int
    L.wp_fit_offset_exponential.argtypes = [c_int, c_void_p, c_void_p, c_void_p, c_void_p]
    _lib = L
    return L


def check(rc):
    if rc != WP_OK:
        msg = lib().wp_last_error().decode(errors="replace")
        raise WhatprotError(f"whatprot_amd error {rc}: {msg}")
