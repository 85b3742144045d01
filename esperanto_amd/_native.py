"""ctypes binding of the in-tree HIP library esperanto_amd/lib/libesp_align.so.

There is no CPU fallback: if the library is missing, or no GPU is visible,
every entry point raises.  Symbol list = include/esp_align.h.
"""
from __future__ import annotations

import ctypes
import os

from .records import CExt, CParams, CTouch

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libesp_align.so")

ESP_OK, ESP_E_INVAL, ESP_E_CAP, ESP_E_DEVICE, ESP_E_NOMEM, ESP_E_BUSY = range(6)
ESP_OPT_FULL_KERNEL, ESP_OPT_TRACE_BUDGET, ESP_OPT_BAND_GRID, ESP_OPT_FULL_PRUNE = 1, 2, 3, 4
ESP_OPT_FULL_PRUNE_SHAPE = 5
ESP_OPT_PLAIN_KERNEL = 6
ESP_OPT_FULL_CENSUS = 7
_STATUS = {1: "invalid argument", 2: "capacity", 3: "device error", 4: "out of memory",
           5: "busy"}

EXPORTS = [
    "esp_create", "esp_destroy", "esp_last_error", "esp_sync", "esp_extend", "esp_extend_hint",
    "esp_band", "esp_myers_best_end", "esp_myers_best_start", "esp_cigar_capacity",
    "esp_gotoh_batch", "esp_myers_batch", "esp_last_kernel_ms", "esp_device_alloc",
    "esp_device_free", "esp_copy_to_device", "esp_copy_to_host", "esp_stitch_cigar",
    "esp_last_timings", "esp_device_copy", "esp_ref_upload", "esp_ref_free",
    "esp_ref_device_bytes", "esp_gather_batch", "esp_mate_rescue_batch", "esp_spliced_batch",
    "esp_last_plan_stats", "esp_server_create", "esp_server_destroy", "esp_server_submit",
    "esp_server_wait", "esp_server_stats_get", "esp_now_ns", "esp_ctx_set_option",
    "esp_unspliced_batch", "esp_window_myers_batch", "esp_window_myers_batch_noalloc",
    "esp_junctions_upload_stranded", "esp_refine_junction_batch", "esp_align_spliced_batch",
    "esp_reextend_batch_run", "esp_intron_chain_batch_run",
    "esp_junctions_upload", "esp_junctions_free",
    "esp_walk_placements", "esp_intron_length_penalty", "esp_fast_lane_batch",
    "esp_engine_myers_batch", "esp_engine_verify_batch", "esp_gotoh_graph_create", "esp_graph_launch",
    "esp_graph_check", "esp_graph_destroy", "esp_autotune_band", "esp_autotune_prune", "esp_myers_infix",
    "esp_mate_rescue_graph_create", "esp_spliced_graph_create",
]


class EspError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"esp status {code} ({_STATUS.get(code, '?')}): {msg}")
        self.code = code


class CBatch(ctypes.Structure):
    _fields_ = [("npairs", ctypes.c_size_t), ("reads", ctypes.c_void_p),
                ("read_off", ctypes.c_void_p), ("read_len", ctypes.c_void_p),
                ("wins", ctypes.c_void_p), ("win_off", ctypes.c_void_p),
                ("win_len", ctypes.c_void_p), ("offset", ctypes.c_void_p),
                ("half", ctypes.c_void_p)]


class CBatchOut(ctypes.Structure):
    _fields_ = [("ext", ctypes.c_void_p), ("touch", ctypes.c_void_p), ("cigar", ctypes.c_void_p),
                ("cig_off", ctypes.c_void_p), ("cig_n", ctypes.c_void_p), ("aux", ctypes.c_void_p)]


class CTiming(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("half", ctypes.c_int32), ("ms", ctypes.c_double)]


class CHostMeta(ctypes.Structure):
    _fields_ = [("read_len", ctypes.c_void_p), ("win_len", ctypes.c_void_p),
                ("offset", ctypes.c_void_p), ("half", ctypes.c_void_p),
                ("max_read_len", ctypes.c_uint32)]


class CPlanStats(ctypes.Structure):
    _fields_ = [("band_pairs", ctypes.c_uint32 * 5), ("strip_pairs", ctypes.c_uint32),
                ("escalated", ctypes.c_uint32), ("full_pruned", ctypes.c_uint32),
                ("full_unpruned", ctypes.c_uint32), ("full_second", ctypes.c_uint32),
                ("full_cells_computed", ctypes.c_uint64), ("full_lane_cells_issued", ctypes.c_uint64),
                ("full_cells_matrix", ctypes.c_uint64), ("full_small_head", ctypes.c_uint32),
                ("full_large_head", ctypes.c_uint32)]


class CChainBatch(ctypes.Structure):
    _fields_ = [("npairs", ctypes.c_size_t), ("reads", ctypes.c_void_p),
                ("read_off", ctypes.c_void_p), ("read_len", ctypes.c_void_p),
                ("contig", ctypes.c_void_p), ("minus", ctypes.c_void_p),
                ("ref_start", ctypes.c_void_p), ("ref_end", ctypes.c_void_p),
                ("read_start", ctypes.c_void_p), ("read_end", ctypes.c_void_p),
                ("allow_dp", ctypes.c_void_p), ("max_read_len", ctypes.c_uint32)]


class CWalkBatch(ctypes.Structure):
    _fields_ = [("nplace", ctypes.c_size_t), ("reads", ctypes.c_void_p),
                ("read_off", ctypes.c_void_p), ("read_len", ctypes.c_void_p),
                ("contig", ctypes.c_void_p), ("pos", ctypes.c_void_p), ("cigar", ctypes.c_void_p),
                ("cig_off", ctypes.c_void_p), ("cig_n", ctypes.c_void_p), ("dir", ctypes.c_void_p)]


class CWalkParams(ctypes.Structure):
    _fields_ = [("known_bonus", ctypes.c_int32), ("gt_ag_bonus", ctypes.c_int32),
                ("gc_ag_bonus", ctypes.c_int32), ("at_ac_bonus", ctypes.c_int32),
                ("noncanonical_penalty", ctypes.c_int32), ("seed_k", ctypes.c_uint32)]


class CFastBatch(ctypes.Structure):
    _fields_ = [("nchain", ctypes.c_size_t), ("reads", ctypes.c_void_p),
                ("read_off", ctypes.c_void_p), ("read_len", ctypes.c_void_p),
                ("contig", ctypes.c_void_p), ("anchors", ctypes.c_void_p),
                ("anc_off", ctypes.c_void_p), ("anc_n", ctypes.c_void_p)]


class CEngineVerify(ctypes.Structure):
    _fields_ = [("n", ctypes.c_size_t), ("reads", ctypes.c_void_p), ("read_off", ctypes.c_void_p),
                ("read_len", ctypes.c_void_p), ("tx", ctypes.c_void_p), ("tx_off", ctypes.c_void_p),
                ("tx_len", ctypes.c_void_p), ("anchor", ctypes.c_void_p)]


class CTuneReport(ctypes.Structure):
    _fields_ = [("half", ctypes.c_int32 * 5), ("occupancy", ctypes.c_int32 * 5),
                ("blocks_per_cu", ctypes.c_int32 * 5), ("ms_default", ctypes.c_double * 5),
                ("ms_best", ctypes.c_double * 5), ("candidates", ctypes.c_int32)]


class CPruneTuneReport(ctypes.Structure):
    _fields_ = [("candidates", ctypes.c_int32), ("half", ctypes.c_int32 * 16), ("pad", ctypes.c_int32 * 16),
                ("ms", ctypes.c_double * 16), ("chosen_half", ctypes.c_int32), ("chosen_pad", ctypes.c_int32),
                ("ms_default", ctypes.c_double), ("ms_best", ctypes.c_double)]


class CMateBatch(ctypes.Structure):
    _fields_ = [("npairs", ctypes.c_size_t), ("reads", ctypes.c_void_p),
                ("read_off", ctypes.c_void_p), ("read_len", ctypes.c_void_p),
                ("mate_contig", ctypes.c_void_p), ("mate_pos", ctypes.c_void_p),
                ("max_read_len", ctypes.c_uint32)]


class CSpliceBatch(ctypes.Structure):
    _fields_ = [("npairs", ctypes.c_size_t), ("reads", ctypes.c_void_p),
                ("read_off", ctypes.c_void_p), ("read_len", ctypes.c_void_p),
                ("contig", ctypes.c_void_p), ("hint", ctypes.c_void_p),
                ("exons", ctypes.c_void_p), ("exon_off", ctypes.c_void_p),
                ("pseudo_off", ctypes.c_void_p), ("pseudo_bytes", ctypes.c_uint64),
                ("n_exons", ctypes.c_size_t)]


class CSpliceCfg(ctypes.Structure):
    """esp_splice_params: SpliceParams (splice.rs:21-53)."""
    _fields_ = [("min_intron", ctypes.c_uint32), ("max_intron", ctypes.c_uint32),
                ("gt_ag_bonus", ctypes.c_int32), ("gc_ag_bonus", ctypes.c_int32), ("at_ac_bonus", ctypes.c_int32),
                ("noncanonical_penalty", ctypes.c_int32), ("known_bonus", ctypes.c_int32),
                ("refine_radius", ctypes.c_uint32)]


class CSplicedChainBatch(ctypes.Structure):
    _fields_ = [("nchain", ctypes.c_size_t), ("reads", ctypes.c_void_p), ("read_off", ctypes.c_void_p),
                ("read_len", ctypes.c_void_p), ("contig", ctypes.c_void_p), ("minus", ctypes.c_void_p),
                ("anchors", ctypes.c_void_p), ("anc_off", ctypes.c_void_p)]


class CReextendBatch(ctypes.Structure):
    _fields_ = [("nplace", ctypes.c_size_t), ("reads", ctypes.c_void_p), ("read_off", ctypes.c_void_p),
                ("read_len", ctypes.c_void_p), ("contig", ctypes.c_void_p), ("pos", ctypes.c_void_p),
                ("cigar", ctypes.c_void_p), ("cig_off", ctypes.c_void_p), ("cig_n", ctypes.c_void_p)]


class CIntronChainBatch(ctypes.Structure):
    _fields_ = [("nread", ctypes.c_size_t), ("reads", ctypes.c_void_p), ("read_off", ctypes.c_void_p),
                ("read_len", ctypes.c_void_p), ("anchors", ctypes.c_void_p), ("anc_off", ctypes.c_void_p),
                ("max_read_len", ctypes.c_uint32)]


class CIntronParams(ctypes.Structure):
    _fields_ = [("min_frag_cov", ctypes.c_uint32), ("refine_window", ctypes.c_uint32),
                ("refine_pattern_len", ctypes.c_uint32), ("intron_max", ctypes.c_uint32)]


_lib = None


def lib():
    """Loads the library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"HIP library not built: {LIB_PATH}; run python -m esperanto_amd.build")
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.POINTER
    vp, sz, u8p, u32p = ctypes.c_void_p, ctypes.c_size_t, P(ctypes.c_uint8), P(ctypes.c_uint32)
    L.esp_create.argtypes = [ctypes.c_int, P(vp)]
    L.esp_destroy.argtypes = [vp]
    L.esp_last_error.argtypes = [vp]
    L.esp_last_error.restype = ctypes.c_char_p
    L.esp_sync.argtypes = [vp]
    L.esp_extend.argtypes = [vp, u8p, sz, u8p, sz, P(CParams), P(CExt), u32p, sz, P(sz)]
    L.esp_extend_hint.argtypes = [vp, u8p, sz, u8p, sz, P(CParams), ctypes.c_int64, P(CExt), u32p,
                                  sz, P(sz)]
    L.esp_band.argtypes = [vp, u8p, sz, u8p, sz, P(CParams), ctypes.c_int, ctypes.c_int64,
                           ctypes.c_int64, P(CExt), P(CTouch), u32p, sz, P(sz)]
    L.esp_myers_best_end.argtypes = [vp, u8p, sz, u8p, sz, P(ctypes.c_int32), P(sz)]
    L.esp_myers_best_start.argtypes = [vp, u8p, sz, u8p, sz, P(ctypes.c_int32), P(sz)]
    L.esp_cigar_capacity.argtypes = [sz]
    L.esp_cigar_capacity.restype = sz
    L.esp_gotoh_batch.argtypes = [vp, P(CBatch), P(CHostMeta), P(CParams), ctypes.c_int,
                                  P(CBatchOut)]
    L.esp_myers_batch.argtypes = [vp, P(CBatch), P(CHostMeta), vp, vp, vp]
    L.esp_last_kernel_ms.argtypes = [vp, P(ctypes.c_double)]
    L.esp_last_timings.argtypes = [vp, P(CTiming), sz, P(sz)]
    L.esp_device_copy.argtypes = [vp, vp, vp, sz]
    L.esp_device_alloc.argtypes = [vp, sz, P(vp)]
    L.esp_device_free.argtypes = [vp, vp]
    L.esp_copy_to_device.argtypes = [vp, vp, vp, sz]
    L.esp_copy_to_host.argtypes = [vp, vp, vp, sz]
    L.esp_stitch_cigar.argtypes = [u32p, sz, ctypes.c_uint32, u32p, sz, u32p, sz, u32p, sz, P(sz)]
    L.esp_last_plan_stats.argtypes = [vp, P(CPlanStats)]
    L.esp_ctx_set_option.argtypes = [vp, ctypes.c_int, ctypes.c_int64]
    L.esp_ref_upload.argtypes = [vp, sz, vp, vp, vp, vp, vp, P(vp)]
    L.esp_ref_free.argtypes = [vp, vp]
    L.esp_ref_device_bytes.argtypes = [vp, P(sz)]
    L.esp_gather_batch.argtypes = [vp, vp, vp, sz, vp, vp]
    L.esp_mate_rescue_batch.argtypes = [vp, vp, P(CMateBatch), P(CParams), vp, vp, vp, vp]
    L.esp_window_myers_batch.argtypes = [vp, vp, vp, sz, vp, ctypes.c_uint32, ctypes.c_uint32,
                                         ctypes.c_int, vp, vp, vp]
    L.esp_window_myers_batch_noalloc.argtypes = L.esp_window_myers_batch.argtypes
    L.esp_unspliced_batch.argtypes = [vp, vp, P(CChainBatch), P(CParams), vp, vp, vp, vp]
    L.esp_spliced_batch.argtypes = [vp, vp, P(CSpliceBatch), P(CParams), P(CBatchOut), vp, vp, vp]
    L.esp_junctions_upload.argtypes = [vp, sz, vp, P(vp)]
    L.esp_junctions_upload_stranded.argtypes = [vp, sz, vp, P(vp)]
    L.esp_refine_junction_batch.argtypes = [vp, vp, vp, vp, sz, vp, P(CSpliceCfg), vp]
    L.esp_align_spliced_batch.argtypes = [vp, vp, vp, P(CSplicedChainBatch), ctypes.c_uint32, P(CSpliceCfg),
                                          P(CParams), P(CBatchOut), vp, vp, vp, vp, vp]
    L.esp_intron_chain_batch_run.argtypes = [vp, vp, vp, P(CIntronChainBatch), ctypes.c_uint32, P(CIntronParams),
                                             P(CParams), P(CBatchOut), vp, vp, vp]
    L.esp_reextend_batch_run.argtypes = [vp, vp, P(CReextendBatch), P(CParams), ctypes.c_int, vp, vp, vp, vp]
    L.esp_junctions_free.argtypes = [vp, vp]
    L.esp_walk_placements.argtypes = [vp, vp, vp, P(CWalkBatch), P(CParams), P(CWalkParams), vp, vp,
                                      vp]
    L.esp_intron_length_penalty.argtypes = [ctypes.c_uint32]
    L.esp_intron_length_penalty.restype = ctypes.c_int32
    L.esp_gotoh_graph_create.argtypes = [vp, P(CBatch), P(CHostMeta), P(CParams), ctypes.c_int, P(CBatchOut),
                                         P(vp)]
    L.esp_graph_launch.argtypes = [vp, vp]
    L.esp_mate_rescue_graph_create.argtypes = [vp, vp, P(CMateBatch), P(CParams), vp, vp, vp, vp, P(vp)]
    L.esp_spliced_graph_create.argtypes = [vp, vp, P(CSpliceBatch), P(CParams), P(CBatchOut), vp, vp, vp,
                                           ctypes.c_uint32, ctypes.c_uint32, P(vp)]
    L.esp_myers_infix.argtypes = [vp, u8p, sz, u8p, sz, P(ctypes.c_int32)]
    L.esp_autotune_band.argtypes = [vp, P(CBatch), P(CHostMeta), P(CParams), P(CBatchOut), ctypes.c_int,
                                    P(CTuneReport)]
    L.esp_autotune_prune.argtypes = [vp, P(CBatch), P(CHostMeta), P(CParams), P(CBatchOut), ctypes.c_int,
                                     P(CPruneTuneReport)]
    L.esp_graph_check.argtypes = [vp, vp]
    L.esp_graph_destroy.argtypes = [vp, vp]
    L.esp_engine_myers_batch.argtypes = [vp, P(CBatch), P(CHostMeta), ctypes.c_int, vp]
    L.esp_engine_verify_batch.argtypes = [vp, P(CEngineVerify), vp, ctypes.c_uint32, vp, vp]
    L.esp_fast_lane_batch.argtypes = [vp, vp, P(CFastBatch), P(CParams), ctypes.c_uint32, ctypes.c_uint32,
                                      vp, vp]
    _lib = L
    return L


def check(rc: int, ctx=None) -> None:
    if rc != ESP_OK:
        msg = lib().esp_last_error(ctx).decode() if ctx else ""
        raise EspError(rc, msg)


class CServerConfig(ctypes.Structure):
    _fields_ = [("params", CParams), ("mem_budget", ctypes.c_uint64), ("workers", ctypes.c_uint32),
                ("max_batch_pairs", ctypes.c_uint32), ("max_pending", ctypes.c_uint32),
                ("max_read_len", ctypes.c_uint32)]


class CRequest(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_uint32), ("npairs", ctypes.c_uint32), ("reads", ctypes.c_void_p),
                ("read_len", ctypes.c_void_p), ("contig", ctypes.c_void_p),
                ("win_start", ctypes.c_void_p), ("win_end", ctypes.c_void_p),
                ("hint", ctypes.c_void_p), ("mate_pos", ctypes.c_void_p)]


class CRequestOut(ctypes.Structure):
    _fields_ = [("ext", ctypes.c_void_p), ("mate", ctypes.c_void_p), ("cigar", ctypes.c_void_p),
                ("cig_n", ctypes.c_void_p)]


class CRequestInfo(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int32), ("done", ctypes.c_int32), ("submit_ns", ctypes.c_uint64),
                ("done_ns", ctypes.c_uint64), ("batch", ctypes.c_uint64)]


class CServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_uint64), ("completed", ctypes.c_uint64),
                ("batches", ctypes.c_uint64), ("pairs", ctypes.c_uint64),
                ("inflight_bytes", ctypes.c_uint64), ("peak_inflight_bytes", ctypes.c_uint64),
                ("pending", ctypes.c_uint32), ("max_batch_requests", ctypes.c_uint32),
                ("stage_ns", ctypes.c_uint64), ("extend_ns", ctypes.c_uint64),
                ("rescue_ns", ctypes.c_uint64), ("scatter_ns", ctypes.c_uint64)]


def server_lib():
    """The library with the serving-loop signatures bound."""
    L = lib()
    if not getattr(L, "_server_bound", False):
        vp, P = ctypes.c_void_p, ctypes.POINTER
        L.esp_server_create.argtypes = [ctypes.c_int, vp, P(CServerConfig), P(vp)]
        L.esp_server_destroy.argtypes = [vp]
        L.esp_server_submit.argtypes = [vp, P(CRequest), P(CRequestOut), P(ctypes.c_uint64)]
        L.esp_server_wait.argtypes = [vp, ctypes.c_uint64, ctypes.c_int64, P(CRequestInfo)]
        L.esp_server_stats_get.argtypes = [vp, P(CServerStats)]
        L.esp_now_ns.argtypes = []
        L.esp_now_ns.restype = ctypes.c_uint64
        L._server_bound = True
    return L
