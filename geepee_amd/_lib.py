"""ctypes binding of libgpx.so (the C-ABI declared in include/gpx.h).

Every entry point takes raw device pointers (FP64, row-major) and a HIP stream and
returns an int status (0 ok, <0 invalid argument / HIP error).  Tensors are torch
CUDA(HIP) tensors used only as device memory; this module never computes anything
itself and raises if the library is missing -- there is no CPU fallback.
"""
import ctypes
import os

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, 'libgpx.so')

_i, _l, _d, _p = ctypes.c_int, ctypes.c_long, ctypes.c_double, ctypes.c_void_p

# name -> (restype, argtypes)
SIGNATURES = {
    'gpx_gemm_workspace': (_l, [_i, _i, _i, _i]),
    'gpx_gemm_batched': (_i, [_i, _i, _i, _i, _i, _d, _p, _i, _l, _p, _i, _l, _d, _p, _i, _l,
                              _i, _i, _p, _i, _p]),
    'gpx_gemm_batched_ex': (_i, [_i, _i, _i, _i, _i, _d, _p, _i, _l, _p, _i, _l, _d, _p, _i, _l,
                                 _d, _p, _i, _l, _i, _i, _p, _i, _p]),
    'gpx_spd_max_m': (_i, []),
    'gpx_spd_inv_batched': (_i, [_i, _i, _p, _i, _l, _p, _i, _l, _p, _p, _p]),
    'gpx_status_fold': (_i, [_i, _p, _p, _p, _p]),
    'gpx_se_kernel': (_i, [_i, _i, _i, _p, _i, _p, _p, _p, _d, _p, _p]),
    'gpx_kfu_bwd_workspace': (_l, [_i, _i, _i]),
    'gpx_kfu_bwd': (_i, [_i, _i, _i, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p]),
    'gpx_dtrace_workspace': (_l, [_i, _i]),
    'gpx_dtrace_kzz': (_i, [_i, _i, _p, _p, _p, _p, _p, _p, _p, _p, _p]),
    'gpx_colsum': (_i, [_i, _i, _p, _d, _d, _p, _p]),
    'gpx_colsum_workspace': (_l, [_i, _i]),
    'gpx_colsum_tall': (_i, [_i, _i, _p, _d, _d, _p, _p, _p]),
    'gpx_eta_to_R': (_i, [_i, _i, _p, _p, _p]),
    'gpx_pack_dR': (_i, [_i, _i, _p, _p, _p, _p]),
    'gpx_prec_form': (_i, [_i, _p, _i, _p, _p, _d, _p, _p]),
    'gpx_kuu_batched': (_i, [_i, _i, _p, _p, _p, _p, _d, _p, _p]),
    'gpx_lincomb': (_i, [_l, _d, _p, _l, _d, _p, _l, _d, _p, _l, _p, _p]),
    'gpx_sym': (_i, [_i, _i, _d, _p, _p, _p]),
    'gpx_det_var': (_i, [_i, _i, _i, _p, _p, _p, _p, _p]),
    'gpx_mm_var': (_i, [_l, _p, _p, _p, _p, _p]),
    'gpx_batched_dot': (_i, [_i, _i, _p, _p, _p, _p]),
    'gpx_gauss_tilted': (_i, [_i, _i, _p, _p, _p, _p, _d, _d, _p, _p, _p, _p, _p]),
    'gpx_gauss_tilted_mc': (_i, [_i, _i, _i, _p, _p, _p, _p, _d, _d, _p, _p, _p, _p]),
    'gpx_probit_tilted': (_i, [_i, _i, _i, _i, _p, _p, _p, _d, _d, _p, _p, _p, _p, _p]),
    'gpx_mc_inputs': (_i, [_i, _i, _i, _p, _p, _p, _p, _p]),
    'gpx_mc_reparam_bwd': (_i, [_i, _i, _i, _p, _p, _p, _p, _p, _p]),
    'gpx_transition_tilted_mc': (_i, [_i, _i, _i, _p, _p, _p, _p, _p, _d, _d, _p, _p, _p, _p, _p, _p]),
    'gpx_transition_tilted': (_i, [_i, _i, _p, _p, _p, _p, _p, _d, _d, _p, _p, _p, _p, _p, _p]),
    'gpx_emission_workspace': (_l, [_i, _i, _i]),
    'gpx_emission_tilted': (_i, [_i, _i, _i, _p, _p, _p, _i, _p, _i, _p, _i, _d, _d, _p, _p, _p,
                                 _p, _p, _p, _p]),
    'gpx_vfe_kl': (_i, [_i, _i, _p, _p, _d, _p, _p, _p, _p]),
    'gpx_gauss_lik_exp': (_i, [_i, _i, _i, _i, _p, _p, _p, _p, _d, _p, _p, _p, _p]),
    'gpx_probit_lik_exp': (_i, [_i, _i, _i, _i, _p, _p, _p, _d, _p, _p, _p, _p]),
    'gpx_transition_lik_exp': (_i, [_i, _i, _i, _i, _p, _p, _p, _p, _p, _d, _p, _p, _p, _p, _p, _p]),
    'gpx_emission_lik_exp_workspace': (_l, [_i, _i]),
    'gpx_emission_lik_exp': (_i, [_i, _i, _i, _p, _p, _p, _i, _p, _i, _p, _i, _d, _p, _p, _p, _p, _p,
                                  _p, _p]),
    'gpx_lvm_x_post': (_i, [_i, _i, _p, _p, _p, _d, _d, _i, _p, _p, _p]),
    'gpx_lvm_vfe_x_grads': (_i, [_i, _i, _p, _p, _p, _p, _p, _d, _d, _i, _d, _p, _p, _p, _p]),
    'gpx_ssm_vfe_x_grads': (_i, [_i, _i, _i, _i, _p, _p, _d, _d, _p, _p, _i, _p, _p, _p, _p, _i, _d,
                                 _p, _p, _p, _p]),
    'gpx_rowscale': (_i, [_i, _i, _i, _p, _p, _p, _p]),
    'gpx_wsum': (_i, [_i, _i, _i, _d, _p, _p, _d, _p, _p]),
    'gpx_dmall': (_i, [_l, _p, _p, _p, _p, _p]),
    'gpx_hyper_grads': (_i, [_i, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p]),
    'gpx_phi': (_i, [_i, _p, _p, _d, _p, _p, _d, _p, _d, _d, _p, _p]),
    'gpx_phi_nat': (_i, [_i, _i, _p, _p, _p, _p, _p, _p, _d, _d, _d, _p, _p]),
    'gpx_lvm_x_cavity': (_i, [_i, _i, _p, _p, _p, _d, _d, _d, _i, _p, _p, _p]),
    'gpx_lvm_x_grads': (_i, [_i, _i, _i, _p, _p, _p, _p, _p, _d, _d, _d, _i, _p, _p, _p, _p]),
    'gpx_ssm_x_cavity': (_i, [_i, _i, _p, _p, _d, _d, _d, _p, _p, _p]),
    'gpx_ssm_x_posterior': (_i, [_i, _i, _p, _p, _d, _d, _p, _p, _p]),
    'gpx_hidden_cavity': (_i, [_i, _i, _p, _p, _d, _p, _p, _p]),
    'gpx_hidden_grads': (_i, [_i, _i, _p, _p, _d, _p, _p, _p, _p, _p, _p]),
    'gpx_ssm_x_grads': (_i, [_i, _i, _i, _i, _p, _p, _d, _d, _d, _p, _p, _i, _p, _p, _p, _p, _i, _p,
                             _p, _p, _p]),
    'gpx_emission_predict': (_i, [_i, _i, _i, _p, _p, _p, _p, _p, _p, _p]),
    'gpx_serve_tile': (_i, []),
    'gpx_serve_desc_sizes': (_i, [_p, _p]),
    'gpx_serve_validate': (_i, [_p]),
    'gpx_predict_grouped': (_i, [_i, _p, _p, _p, _p, _p, _i, _p]),
    'gpx_sym_pairmajor': (_i, [_i, _i, _p, _p, _p]),
    'gpx_adam_step': (_i, [_l, _p, _p, _p, _p, _d, _d, _d, _d, _i, _p, _p, _p]),
    'gpx_fused_fwd_workspace': (_l, [_i, _i, _i, _i]),
    'gpx_fused_mm_fwd': (_i, [_i, _i, _i, _i, _p, _p, _p, _p, _p, _p, _p, _i, _i, _p, _p, _p, _p, _p]),
    'gpx_fused_fwd_noalloc_max': (_l, []),
    'gpx_fused_mm_fwd_noalloc': (_i, [_i, _i, _i, _i, _p, _p, _p, _p, _p, _p, _p, _i, _i, _p, _p, _p, _p]),
    'gpx_fused_bwd_workspace': (_l, [_i, _i, _i, _i, _i, _i]),
    'gpx_fused_mm_bwd': (_i, [_i, _i, _i, _i, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _i, _i, _p, _p, _p,
                              _p, _p, _p, _p, _p]),
    'gpx_fused_mm_bwd_main': (_i, [_i, _i, _i, _i, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _i, _i, _p, _p,
                                   _p, _p, _p]),
    'gpx_fused_mm_bwd_points': (_i, [_i, _i, _i, _i, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _i, _i, _p, _p,
                                     _p, _p]),
    'gpx_fused_mm_bwd_pairs': (_i, [_i, _i, _i, _i, _p, _p, _p, _p, _p, _i, _i, _p, _p, _p]),
    'gpx_fused_mm_bwd_tail': (_i, [_i, _i, _i, _i, _p, _p, _p, _i, _i, _p, _p, _p, _p, _p]),
    'gpx_psi_bwd_tail': (_i, [_i, _i, _i, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p,
                              _p, _p]),
    'gpx_psi_qmax': (_i, []),
    'gpx_psi_fwd': (_i, [_i, _i, _i, _p, _p, _p, _p, _p, _p, _p, _p]),
    'gpx_psi_bwd_workspace': (_l, [_i, _i, _i]),
    'gpx_psi_bwd': (_i, [_i, _i, _i, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p,
                         _p, _p]),
}


GPX_EINVAL, GPX_EHIP, GPX_ETOOBIG = -1, -2, -3


class GpxError(RuntimeError):
    pass


_lib = None


def load():
    """Load libgpx.so and declare every entry point (works without a GPU)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise GpxError('libgpx.so is not built (run python -m geepee_amd.build_lib)')
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


_probe = None


def set_probe(p):
    """Install (or clear with None) a Probe that brackets one entry point with HIP
    events on the launch stream."""
    global _probe
    _probe = p


class Probe(object):
    """Times every launch of entry point ``target`` with a pair of HIP events recorded on
    the stream the kernels are launched on."""

    def __init__(self, target):
        self.target = target
        self.pairs = []

    def count(self):
        return len(self.pairs)

    def mean_ms(self):
        torch.cuda.synchronize()
        if not self.pairs:
            return float('nan')
        return sum(a.elapsed_time(b) for a, b in self.pairs) / len(self.pairs)


def call(name, *args):
    p = _probe
    if p is not None and name == p.target:
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        rc = getattr(load(), name)(*args)
        b.record()
        p.pairs.append((a, b))
    else:
        rc = getattr(load(), name)(*args)
    if rc != 0:
        if rc == GPX_ETOOBIG:
            raise GpxError('%s: the shape needs more workspace than the allocation-free '
                           'entry point\'s fixed scratch holds (status %d)' % (name, rc))
        raise GpxError('%s failed with status %d' % (name, rc))
    return rc


def query(name, *args):
    return getattr(load(), name)(*args)


def P(t):
    """Device pointer of a float64 CUDA tensor (views allowed), or None."""
    if t is None:
        return None
    if not t.is_cuda:
        raise GpxError('expected a device tensor')
    return t.data_ptr()


def S():
    return torch.cuda.current_stream().cuda_stream


def require_device():
    if not torch.cuda.is_available():
        raise GpxError('no HIP device: the geepee_amd compute path runs only on the GPU')
    load()


# ---------------------------------------------------------------------------
# thin typed wrappers
# ---------------------------------------------------------------------------

def gemm(tA, tB, m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch=1,
         splits=1, work=None, tile=0):
    call('gpx_gemm_batched', int(tA), int(tB), m, n, k, float(alpha), P(A), lda, sA, P(B), ldb, sB,
         float(beta), P(C), ldc, sC, batch, splits, P(work), tile, S())


def mm(A, B, C, tA=False, tB=False, alpha=1.0, beta=0.0, splits=1, work=None, tile=0,
       E=None, gamma=0.0):
    """C = alpha op(A) op(B) + beta C + gamma E for 2-D or batched 3-D contiguous
    tensors; a 2-D operand (E included) is broadcast over the batch of a 3-D one."""
    batch = max(t.shape[0] if t.dim() == 3 else 1 for t in (A, B, C))
    ar, ac = A.shape[-2:]
    br, bc = B.shape[-2:]
    m, k = (ac, ar) if tA else (ar, ac)
    k2, n = (bc, br) if tB else (br, bc)
    assert k == k2, (A.shape, B.shape, tA, tB)
    assert C.shape[-2:] == (m, n), (C.shape, m, n)
    for t in (A, B, C):
        assert t.is_contiguous() and t.dtype == torch.float64
    sA = ar * ac if A.dim() == 3 else 0
    sB = br * bc if B.dim() == 3 else 0
    sC = m * n if C.dim() == 3 else 0
    if splits > 1 and work is None:
        work = torch.empty(splits * batch * m * n, dtype=torch.float64, device=C.device)
    if E is None:
        gemm(tA, tB, m, n, k, alpha, A, ac, sA, B, bc, sB, beta, C, n, sC, batch, splits, work, tile)
        return C
    assert E.is_contiguous() and E.dtype == torch.float64 and E.shape[-2:] == (m, n), E.shape
    sE = m * n if E.dim() == 3 else 0
    call('gpx_gemm_batched_ex', int(tA), int(tB), m, n, k, float(alpha), P(A), ac, sA, P(B), bc, sB,
         float(beta), P(C), n, sC, float(gamma), P(E), n, sE, batch, splits, P(work), tile, S())
    return C


def spd_inv(A, out=None, logdet=None, info=None):
    """Batched SPD inverse / log-determinant of A (b, M, M) or (M, M)."""
    M = A.shape[-1]
    batch = A.shape[0] if A.dim() == 3 else 1
    call('gpx_spd_inv_batched', M, batch, P(A), M, M * M, P(out), M, M * M, P(logdet), P(info), S())
    return out
