"""Gradients of the solute field through a recorded CA trajectory.

A forward run of a `Subblock` records which branch of fs_change_diffuse
(growth.c:221-273) every cell took at every step (`record`).  With those
branches frozen -- and with them every nucleation and capture decision --
the final concentration field is a smooth function of the diffusivities
D_l, D_s, the partition coefficient k and the initial concentration field.
`SoluteMap` is that function as a `torch.autograd.Function`: its forward
replays the steps on the GPU (`mmca_replay_step`), its backward runs the
hand-written adjoint kernels (`mmca_adjoint_step`) backwards in time with
sqrt(T) checkpointing: the forward keeps every K-th state (K = ceil(sqrt(T))),
the backward recomputes one K-step segment at a time from its checkpoint.

All state tensors are FP64, on the GPU, flattened padded boxes of
(nx+2)(ny+2)(nz+2) cells in the reference's layout.  There is no CPU path.
"""
import ctypes as C
import math

import numpy as np
import torch

from . import abi
from .lib import _check, lib


def _params_with(params, d_liq, d_sol, k):
    p = abi.Params()
    C.pointer(p)[0] = params
    p.d_liq = float(d_liq)
    p.d_sol = float(d_sol)
    p.part_coef = float(k)
    p.precision = abi.PREC_FP64
    return p


def _ptr(t):
    return C.c_void_p(t.data_ptr())


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def record_bytes(params):
    """Bytes of one step's branch record (3 bits per cell, mmca_record_bytes)."""
    n = C.c_int64(0)
    _check(lib().mmca_record_bytes(C.byref(params), C.byref(n)), "mmca_record_bytes")
    return n.value


def unpack_codes(params, packed):
    """Branch codes per padded cell, uint8 (nsteps, nz+2, ny+2, nx+2), from a
    packed record (nsteps, record_bytes) on any device: interior cells get
    their MMCA_BRANCH_* code, halo cells 0."""
    nx, ny, nz = params.nx, params.ny, params.nz
    w64 = (nx + 63) // 64
    T = packed.shape[0]
    words = packed.contiguous().view(torch.int64).view(T, nz + 2, ny + 2, w64, 3)
    bit = torch.arange(64, device=packed.device, dtype=torch.int64)
    out = torch.zeros((T, nz + 2, ny + 2, w64 * 64), dtype=torch.uint8, device=packed.device)
    for p in range(3):
        b = (words[..., p].unsqueeze(-1) >> bit) & 1
        out |= (b.reshape(T, nz + 2, ny + 2, w64 * 64) << p).to(torch.uint8)
    full = torch.zeros((T, nz + 2, ny + 2, nx + 2), dtype=torch.uint8, device=packed.device)
    full[:, 1:-1, 1:-1, 1:-1] = out[:, 1:-1, 1:-1, :nx]
    return full


class Trajectory:
    """Branch record of `nsteps` steps after timestep `t0`, with the state
    (ce, cl, fs) at t0 and the parameters of the recorded run."""

    def __init__(self, params, t0, codes, ce0, cl0, fs0):
        self.params = params
        self.t0 = t0
        self.codes = codes            # uint8 (nsteps, record_bytes), cuda: 3 bits per cell
        self.ce0, self.cl0, self.fs0 = ce0, cl0, fs0
        self.shape = (params.nz + 2, params.ny + 2, params.nx + 2)

    @property
    def nsteps(self):
        return self.codes.shape[0]

    @property
    def padded_cells(self):
        return self.shape[0] * self.shape[1] * self.shape[2]

    def unpacked_codes(self):
        """The record as one uint8 code per padded cell (nsteps, P)."""
        return unpack_codes(self.params, self.codes).view(self.nsteps, -1)

    def interior(self, field):
        """Interior view (nz, ny, nx) of a flattened padded field."""
        return field.view(self.shape)[1:-1, 1:-1, 1:-1]


def record(sb, nsteps):
    """Advance `sb` (a FP64 Subblock) by nsteps while recording the branches."""
    dev = torch.device("cuda", torch.cuda.current_device())
    ce0 = torch.from_numpy(sb.download(abi.FIELD_CE).reshape(-1)).to(dev)
    cl0 = torch.from_numpy(sb.download(abi.FIELD_CL).reshape(-1)).to(dev)
    fs0 = torch.from_numpy(sb.download(abi.FIELD_FS).reshape(-1)).to(dev)
    codes = torch.zeros((nsteps, record_bytes(sb.params)), dtype=torch.uint8, device=dev)
    t0 = sb.timestep
    torch.cuda.synchronize()          # the library writes on its own stream
    sb.set_record(codes.data_ptr(), nsteps)
    try:
        sb.step(nsteps)
    finally:
        sb.set_record(None, 0)
    p = abi.Params()
    C.pointer(p)[0] = sb.params
    return Trajectory(p, t0, codes, ce0, cl0, fs0)


def replay_step(params, timestep, codes, state, out):
    """out = one frozen-branch step of state; both are (ce, cl, fs) tuples."""
    _check(lib().mmca_replay_step(C.byref(params), timestep, _ptr(codes), _ptr(state[0]), _ptr(state[1]),
                                  _ptr(state[2]), _ptr(out[0]), _ptr(out[1]), _ptr(out[2]), _stream()),
           "mmca_replay_step")


def replay(traj, d_liq, d_sol, k, ce0=None, keep_every=0):
    """Final state of the frozen-branch map; with keep_every > 0 also the
    states at steps 0, K, 2K, ... (checkpoints)."""
    p = _params_with(traj.params, d_liq, d_sol, k)
    cur = [(traj.ce0 if ce0 is None else ce0).clone(), traj.cl0.clone(), traj.fs0.clone()]
    nxt = [t.clone() for t in cur]
    keep = []
    for s in range(traj.nsteps):
        if keep_every and s % keep_every == 0:
            keep.append(tuple(t.clone() for t in cur))
        replay_step(p, traj.t0 + s + 1, traj.codes[s], cur, nxt)
        cur, nxt = nxt, cur
    return tuple(cur), keep


def checkpoint_interval(nsteps):
    return max(1, int(math.ceil(math.sqrt(nsteps))))


def backward(traj, params, checkpoints, K, g_ce, g_cl=None, g_fs=None):
    """Adjoint sweep: derivatives of the loss with respect to the state at t0
    and to (d_liq, d_sol, k), given its derivatives with respect to the final
    state.  `checkpoints[i]` is the state after i*K steps."""
    L = lib()
    nb = C.c_int64(0)
    _check(L.mmca_adjoint_workspace_bytes(C.byref(params), C.byref(nb)), "mmca_adjoint_workspace_bytes")
    dev = g_ce.device
    ws = torch.zeros((nb.value + 7) // 8, dtype=torch.float64, device=dev)
    g = [g_ce.clone(),
         torch.zeros_like(g_ce) if g_cl is None else g_cl.clone(),
         torch.zeros_like(g_ce) if g_fs is None else g_fs.clone()]
    g_param = torch.zeros(3, dtype=torch.float64, device=dev)
    T = traj.nsteps
    nseg = (T + K - 1) // K
    for seg in range(nseg - 1, -1, -1):
        s0 = seg * K
        s1 = min(T, s0 + K)
        states = [checkpoints[seg]]
        for s in range(s0, s1 - 1):
            out = tuple(torch.empty_like(t) for t in states[-1])
            for o, i in zip(out, states[-1]):
                o.copy_(i)                      # halo cells are never written
            replay_step(params, traj.t0 + s + 1, traj.codes[s], states[-1], out)
            states.append(out)
        for s in range(s1 - 1, s0 - 1, -1):
            st = states[s - s0]
            _check(L.mmca_adjoint_step(C.byref(params), traj.t0 + s + 1, _ptr(traj.codes[s]), _ptr(st[0]),
                                       _ptr(st[1]), _ptr(st[2]), _ptr(g[0]), _ptr(g[1]), _ptr(g[2]), _ptr(g[0]),
                                       _ptr(g[1]), _ptr(g[2]), _ptr(g_param), _ptr(ws), _stream()),
                   "mmca_adjoint_step")
        del states
    return g, g_param


class SoluteMap(torch.autograd.Function):
    """ce_T = F(d_liq, d_sol, k, ce0) along a recorded trajectory.

    d_liq, d_sol, k: 0-dim float64 tensors; ce0: flattened padded FP64 field
    on the GPU.  Returns the flattened padded ce after traj.nsteps steps."""

    @staticmethod
    def forward(ctx, d_liq, d_sol, k, ce0, traj):
        dl, ds, kk = float(d_liq), float(d_sol), float(k)
        K = checkpoint_interval(traj.nsteps)
        (ce, cl, fs), keep = replay(traj, dl, ds, kk, ce0=ce0.detach(), keep_every=K)
        ctx.traj = traj
        ctx.K = K
        ctx.keep = keep
        ctx.pvals = (dl, ds, kk)
        ctx.shapes = (d_liq, d_sol, k)
        return ce

    @staticmethod
    def backward(ctx, g_ce):
        traj = ctx.traj
        params = _params_with(traj.params, *ctx.pvals)
        g, g_param = backward(traj, params, ctx.keep, ctx.K, g_ce.contiguous())
        d_liq, d_sol, k = ctx.shapes
        gp = [g_param[i].to(t.device).reshape(t.shape) for i, t in enumerate((d_liq, d_sol, k))]
        return gp[0], gp[1], gp[2], g[0], None


def solute_map(traj, d_liq, d_sol, k, ce0=None):
    """Differentiable final ce of the recorded trajectory (see SoluteMap)."""
    if ce0 is None:
        ce0 = traj.ce0
    return SoluteMap.apply(d_liq, d_sol, k, ce0, traj)


def l2_loss(traj, ce_T, target):
    """sum over the interior of (ce_T - target)^2 (flattened padded fields)."""
    d = traj.interior(ce_T) - traj.interior(target)
    return (d * d).sum()


def l2_loss_difference(traj, ce_a, ce_b, target):
    """l2_loss(ce_a) - l2_loss(ce_b), formed without cancellation: the sum
    over the interior of (a - b)(a + b - 2 t), each term from the difference
    of two nearby values, summed exactly (math.fsum).  For a central finite
    difference of a weakly coupled parameter the two losses agree to many
    digits, and their difference as two rounded sums would be mostly noise."""
    a = traj.interior(ce_a).double()
    b = traj.interior(ce_b).double()
    t = traj.interior(target).double()
    terms = ((a - b) * (a + b - 2.0 * t)).reshape(-1).cpu().numpy()
    return math.fsum(terms.tolist())
