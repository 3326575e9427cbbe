"""Latent-variable (x-factor) plumbing of SGPLVM and SGPSSM on the device."""
import torch

from . import dops
from ._lib import P, S, call

F64 = torch.float64


def lvm_x_state(x1, x2, idx, alpha, prior_mean, prior_var, nat_param):
    """Cavity inputs of the minibatch rows idx (an int64 device tensor, or None for all
    rows) (aep_models.py:840-861)."""
    N, Q = x1.shape
    B = N if idx is None else idx.numel()
    mcav = torch.empty(B, Q, dtype=F64, device=x1.device)
    vcav = torch.empty(B, Q, dtype=F64, device=x1.device)
    call('gpx_lvm_x_cavity', B, Q, P(idx), P(x1), P(x2), float(alpha), float(prior_mean),
         float(prior_var), int(bool(nat_param)), P(mcav), P(vcav), S())
    return {'mcav': mcav, 'vcav': vcav, 'idx': idx, 'x1': x1, 'x2': x2, 'B': B}


def lvm_x_grads(xs, gmx, gvx, N, B, alpha, prior_mean, prior_var, Din, nat_param):
    """Gradients w.r.t. x1, x2 and the x log-partition energy term
    (aep_models.py:784-805, 817-838; base_models.py:913-929)."""
    x1, x2, idx = xs['x1'], xs['x2'], xs['idx']
    g1 = torch.zeros_like(x1) if idx is not None else torch.empty_like(x1)
    g2 = torch.zeros_like(x2) if idx is not None else torch.empty_like(x2)
    scal = torch.empty(2 + 2 * 64, dtype=F64, device=x1.device)
    call('gpx_lvm_x_grads', N, B, Din, P(idx), P(x1), P(x2), P(gmx.contiguous()), P(gvx.contiguous()),
         float(alpha), float(prior_mean), float(prior_var), int(bool(nat_param)), P(g1), P(g2),
         P(scal), S())
    return g1, g2, scal


def ssm_cavity(xf1, x2, alpha, prior_mean, prior_var):
    N, Q = xf1.shape
    cm = torch.empty_like(xf1)
    cv = torch.empty_like(xf1)
    call('gpx_ssm_x_cavity', N, Q, P(xf1), P(x2), float(alpha), float(prior_mean), float(prior_var),
         P(cm), P(cv), S())
    return cm, cv


def ssm_grads(xf1, x2, alpha, prior_mean, prior_var, s, e, dup, vup, dmt, dvt, dmn, dvn):
    N, Q = xf1.shape
    g1 = torch.empty_like(xf1)
    g2 = torch.empty_like(xf1)
    scal = torch.empty(2 + 2 * 64, dtype=F64, device=xf1.device)
    call('gpx_ssm_x_grads', N, Q, s, e, P(xf1), P(x2), float(alpha), float(prior_mean),
         float(prior_var), P(dup), P(vup), dup.stride(0), P(dmt), P(dvt), P(dmn), P(dvn),
         dmn.stride(0), P(g1), P(g2), P(scal), S())
    return g1, g2, scal
