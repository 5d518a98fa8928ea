"""Warm-start network on the GPU: DeepSets encoders + per-point MLP.

Per problem (sources x: n x 16 with weights a, targets y: m x 16 with
weights b, reg), with points scaled by 1/sqrt(max cost) so every problem is
scale free:

    e_j = [y_j, m b_j - 1, lr]       H1 = act(e W_e1 + b_e1)
    H2  = act(H1 W_e2 + b_e2)        zt = mean_j H2      (target encoder)
    s_i = [x_i, n a_i - 1, lr]       S1 = act(s W_s1 + b_s1)
    S2  = act(S1 W_s2 + b_s2)        zs = mean_i S2      (source encoder)
    c   = zt W_m1z + zs W_m1s + b_m1 G1 = act(s W_m1x + c)
    G2  = act(G1 W_m2 + b_m2)        f  = |x_i|^2 + G2 w_m3 + b_m3

with act the leaky ReLU (slope 0.01) and lr = (log reg + 4.147) / 1.151.

``[s_i, zt, zs] W_m1`` is split so the pooled terms are computed once per
problem.  The exact |x_i|^2 term leaves the network the c-transform part of
the potential (a soft-max of functions affine in x).  f is the dual
potential in units of the normalised cost; the solver starts from
u0 = f / reg.

Every forward/backward step runs in libpot_amd.so: the four wide layers on
the bf16 MFMA GEMM (fp32 accumulation, fused bias / per-problem broadcast /
ReLU / ReLU-mask epilogues), the rest on small fp32 kernels; weights are
fp32 masters updated by the library's Adam kernel, with bf16 copies for the
GEMMs.  Inference runs each two-layer chain (and its segment mean or head)
as one fused kernel whose hidden activations never reach HBM.  The float64 NumPy restatement in oracle/warmnet.py is the checker.
"""

import ctypes

import numpy as np
import torch

from .. import _device, _lib

H = 256
DX = 16
KIN = 32  # input width padded to one GEMM K slice

# op codes of potamd_mlp_op
OP_BUILD, OP_SEGMEAN, OP_VECMAT, OP_VECMATT, OP_HEAD, OP_LOSS = 0, 1, 2, 3, 4, 5
OP_HEADBACK, OP_HEADWGRAD, OP_COLSUM, OP_SEGSUM, OP_POOLBACK, OP_OUTER = 6, 7, 8, 9, 10, 11
OP_ADAM, OP_TOBF16, OP_ROWNORM = 12, 13, 14

PARAM_SHAPES = {
    "W_e1": (KIN, H), "b_e1": (H,), "W_e2": (H, H), "b_e2": (H,),
    "W_s1": (KIN, H), "b_s1": (H,), "W_s2": (H, H), "b_s2": (H,),
    "W_m1x": (KIN, H), "W_m1z": (H, H), "W_m1s": (H, H), "b_m1": (H,),
    "W_m2": (H, H), "b_m2": (H,), "w_m3": (H,), "b_m3": (1,),
}
GEMM_WEIGHTS = ("W_e1", "W_e2", "W_s1", "W_s2", "W_m1x", "W_m2")
PADDED_INPUT = ("W_e1", "W_s1", "W_m1x")  # rows >= DX + 2 multiply zero padding
R0 = 1e-2


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


class WarmStartNet:
    def __init__(self, params, device):
        self.device = device
        self.w = {k: torch.as_tensor(np.asarray(v, dtype=np.float32)).to(device).contiguous()
                  for k, v in params.items()}
        self.wb = {k: torch.empty(self.w[k].shape, dtype=torch.bfloat16, device=device)
                   for k in GEMM_WEIGHTS}
        self.g = {k: torch.zeros_like(v) for k, v in self.w.items()}
        self.m1 = {k: torch.zeros_like(v) for k, v in self.w.items()}
        self.m2 = {k: torch.zeros_like(v) for k, v in self.w.items()}
        self.step_count = 0
        self.lib = _lib.load()
        self.sync_bf16()

    # ------------------------------------------------------------------ io
    @staticmethod
    def init_params(seed=0):
        """He-initialised parameters (padded input rows are zero)."""
        rng = np.random.default_rng(seed)

        def lin(fan_in, rows, cols):
            return rng.standard_normal((rows, cols)) * np.sqrt(2.0 / fan_in)
        p = {k: np.zeros(s) for k, s in PARAM_SHAPES.items()}
        p["W_e1"][:DX + 2] = lin(DX + 2, DX + 2, H)
        p["W_e2"] = lin(H, H, H)
        p["W_s1"][:DX + 2] = lin(DX + 2, DX + 2, H)
        p["W_s2"] = lin(H, H, H)
        p["W_m1x"][:DX + 2] = lin(DX + 2 + 2 * H, DX + 2, H)
        p["W_m1z"] = lin(DX + 2 + 2 * H, H, H)
        p["W_m1s"] = lin(DX + 2 + 2 * H, H, H)
        p["W_m2"] = lin(H, H, H)
        # zero head: training starts from the exact |x|^2 predictor
        return p

    @classmethod
    def load(cls, path, device):
        with np.load(path, allow_pickle=False) as z:
            params = {k: z[k] for k in PARAM_SHAPES}
        return cls(params, device)

    def save(self, path, **meta):
        arrs = {k: v.detach().cpu().numpy() for k, v in self.w.items()}
        arrs.update({k: np.asarray(v) for k, v in meta.items()})
        np.savez(path, **arrs)

    def params_numpy(self):
        return {k: v.detach().cpu().double().numpy() for k, v in self.w.items()}

    # ------------------------------------------------------------- kernels
    def _stream(self):
        return ctypes.c_void_p(_device.stream_ptr(self.device))

    def _op(self, op, rows=0, seg_len=1, P=0, k=0, n=0, count=0, flag=0, scalar=0.0, hp=None,
            x=None, y=None, z=None, w=None, out=None, out2=None):
        hp_arr = None
        if hp is not None:
            hp_arr = (ctypes.c_float * 7)(*hp)
        rc = self.lib.potamd_mlp_op(op, rows, seg_len, P, k, n, count, flag, float(scalar),
                                    hp_arr, _p(x), _p(y), _p(z), _p(w), _p(out), _p(out2),
                                    self._stream())
        _lib.check(rc, f"potamd_mlp_op({op})")

    def _gemm(self, A, B, C, M, N, K, *, transA=0, transB=0, out_kind=0, bias=None,
              rowvec=None, seg_len=1, relu=0, mask=None, splits=1, lda=None, ldb=None):
        lda = lda or (M if transA else K)
        ldb = ldb or (K if transB else N)
        rc = self.lib.potamd_gemm_bf16(M, N, K, _p(A), lda, transA, _p(B), ldb, transB, _p(C), N,
                                       out_kind, 1.0, _p(bias), _p(rowvec), seg_len, relu,
                                       _p(mask), N, splits, self._stream())
        _lib.check(rc, "potamd_gemm_bf16")

    def sync_bf16(self):
        for k in GEMM_WEIGHTS:
            self._op(OP_TOBF16, count=self.w[k].numel(), x=self.w[k], out=self.wb[k])

    # ------------------------------------------------------------- inputs
    def build_inputs(self, pts, P, seg_len, w, regs, cmax):
        """(P*seg_len, KIN) bf16 rows [scaled point, seg_len * weight, log reg, 0...]."""
        rows = P * seg_len
        out = torch.empty((rows, KIN), dtype=torch.bfloat16, device=self.device)
        self._op(OP_BUILD, rows=rows, seg_len=seg_len, n=DX, k=KIN, flag=1,
                 x=pts.contiguous(), y=w, z=regs, w=cmax, out=out)
        return out

    def row_norms(self, pts, P, seg_len, cmax):
        """(P*seg_len,) fp32 squared norms of the scaled points."""
        out = torch.empty(P * seg_len, dtype=torch.float32, device=self.device)
        self._op(OP_ROWNORM, rows=P * seg_len, seg_len=seg_len, n=DX, x=pts.contiguous(),
                 w=cmax, out=out)
        return out

    # ------------------------------------------------------------ forward
    def _encode(self, inp, rows, P, seg, W1, b1, W2, b2):
        dev = self.device
        A = torch.empty((rows, H), dtype=torch.bfloat16, device=dev)
        self._gemm(inp, self.wb[W1], A, rows, H, KIN, bias=self.w[b1], relu=1)
        B = torch.empty_like(A)
        self._gemm(A, self.wb[W2], B, rows, H, H, bias=self.w[b2], relu=1)
        z = torch.empty((P, H), dtype=torch.float32, device=dev)
        self._op(OP_SEGMEAN, seg_len=seg, P=P, k=H, x=B, out=z)
        return A, B, z

    def _fused(self, mode, inp, rows, seg, W1, bias1, W2, b2, out, w3=None, b3=None, qn=None,
               pts=None, wts=None, regs=None, cmax=None):
        rc = self.lib.potamd_mlp_fused(mode, rows, seg, _p(inp), _p(pts), _p(wts), _p(regs),
                                       _p(cmax), _p(self.wb[W1]), _p(bias1), _p(self.wb[W2]),
                                       _p(self.w[b2]), _p(w3), _p(b3), _p(qn), _p(out),
                                       self._stream())
        _lib.check(rc, "potamd_mlp_fused")

    # Points per cloud from which forward_points' encoders estimate their
    # segment means (0: every point).  On the 1024-point family, 256 points
    # cost the start 0.2 iterations (36.04 against 35.85 mean) and cut the
    # encoders from 2 x 0.65 to about 2 x 0.2 ms per 4096 problems.
    ENC_POINTS = 256

    def _enc_sample(self, Z, w, P, n):
        """The points of each cloud the encoder reads, with their per-point
        weights: all of them when n <= ENC_POINTS, else ENC_POINTS points at
        stride n // ENC_POINTS, the weights scaled by n / ENC_POINTS so that
        the input feature n w - 1 keeps its value."""
        k = int(self.ENC_POINTS)
        if k <= 0 or n <= k:
            return Z, w, n
        s = n // k
        Zs = Z.view(P, n, DX)[:, :s * k:s].reshape(-1, DX).contiguous()
        ws = (w.view(P, n)[:, :s * k:s] * (n / k)).reshape(-1).contiguous()
        return Zs, ws, k

    def forward_points(self, X, Y, a, b, regs, cmax, P, n, m):
        """Inference forward straight from the points (n, m multiples of 64): the
        fused kernels build the input rows themselves.  X: (P*n, 16) fp32,
        a: (P*n,) fp64, regs: (P,) fp64, cmax: (P,) fp32.  The encoders' means
        come from ``ENC_POINTS`` points of a larger cloud (``_enc_sample``);
        with every point it is ``forward`` up to fp32 rounding order."""
        dev = self.device
        w = self.w
        X = X.reshape(-1, DX).float().contiguous()
        Y = Y.reshape(-1, DX).float().contiguous()
        Ye, be, me = self._enc_sample(Y, b, P, m)
        zt = torch.empty((P, H), dtype=torch.float32, device=dev)
        self._fused(0, None, P * me, me, "W_e1", w["b_e1"], "W_e2", "b_e2", zt, pts=Ye,
                    wts=be, regs=regs, cmax=cmax)
        Xe, ae, ne = self._enc_sample(X, a, P, n)
        zs = torch.empty((P, H), dtype=torch.float32, device=dev)
        self._fused(0, None, P * ne, ne, "W_s1", w["b_s1"], "W_s2", "b_s2", zs, pts=Xe,
                    wts=ae, regs=regs, cmax=cmax)
        c = torch.empty((P, H), dtype=torch.float32, device=dev)
        self._op(OP_VECMAT, P=P, k=H, n=H, x=zt, w=w["W_m1z"], y=w["b_m1"], out=c)
        c2 = torch.empty_like(c)
        self._op(OP_VECMAT, P=P, k=H, n=H, flag=1, x=zs, w=w["W_m1s"], y=c, out=c2)
        qn = self.row_norms(X, P, n, cmax)
        f = torch.empty(P * n, dtype=torch.float32, device=dev)
        self._fused(1, None, P * n, n, "W_m1x", c2, "W_m2", "b_m2", f, w3=w["w_m3"],
                    b3=w["b_m3"], qn=qn, pts=X, wts=a, regs=regs, cmax=cmax)
        return f

    def forward_fused(self, Sin, Ein, qn, P, n, m):
        """Inference forward with the fused two-layer kernel (n, m multiples of
        64): the hidden activations stay on chip; same result as ``forward``
        up to fp32 rounding order and the head's bf16 hi + lo split of w3."""
        dev = self.device
        w = self.w
        zt = torch.empty((P, H), dtype=torch.float32, device=dev)
        self._fused(0, Ein, P * m, m, "W_e1", w["b_e1"], "W_e2", "b_e2", zt)
        zs = torch.empty((P, H), dtype=torch.float32, device=dev)
        self._fused(0, Sin, P * n, n, "W_s1", w["b_s1"], "W_s2", "b_s2", zs)
        c = torch.empty((P, H), dtype=torch.float32, device=dev)
        self._op(OP_VECMAT, P=P, k=H, n=H, x=zt, w=w["W_m1z"], y=w["b_m1"], out=c)
        c2 = torch.empty_like(c)
        self._op(OP_VECMAT, P=P, k=H, n=H, flag=1, x=zs, w=w["W_m1s"], y=c, out=c2)
        f = torch.empty(P * n, dtype=torch.float32, device=dev)
        self._fused(1, Sin, P * n, n, "W_m1x", c2, "W_m2", "b_m2", f, w3=w["w_m3"],
                    b3=w["b_m3"], qn=qn)
        return f

    def forward(self, Sin, Ein, qn, P, n, m, keep=False):
        """f (P*n,) from source rows Sin, target rows Ein and source norms qn.

        Without ``keep`` (inference) and with n, m multiples of 64 this is the
        fused kernel; training keeps the per-layer GEMM path, whose
        activations the backward pass reads."""
        if not keep and n % 64 == 0 and m % 64 == 0:
            return self.forward_fused(Sin, Ein, qn, P, n, m)
        dev = self.device
        w, wb = self.w, self.wb
        H1, H2, zt = self._encode(Ein, P * m, P, m, "W_e1", "b_e1", "W_e2", "b_e2")
        S1, S2, zs = self._encode(Sin, P * n, P, n, "W_s1", "b_s1", "W_s2", "b_s2")
        c = torch.empty((P, H), dtype=torch.float32, device=dev)
        self._op(OP_VECMAT, P=P, k=H, n=H, x=zt, w=w["W_m1z"], y=w["b_m1"], out=c)
        c2 = torch.empty_like(c)
        self._op(OP_VECMAT, P=P, k=H, n=H, flag=1, x=zs, w=w["W_m1s"], y=c, out=c2)
        G1 = torch.empty((P * n, H), dtype=torch.bfloat16, device=dev)
        self._gemm(Sin, wb["W_m1x"], G1, P * n, H, KIN, rowvec=c2, seg_len=n, relu=1)
        G2 = torch.empty_like(G1)
        self._gemm(G1, wb["W_m2"], G2, P * n, H, H, bias=w["b_m2"], relu=1)
        f = torch.empty(P * n, dtype=torch.float32, device=dev)
        self._op(OP_HEAD, rows=P * n, k=H, x=G2, w=w["w_m3"], y=w["b_m3"], z=qn, out=f)
        if keep:
            self._cache = (Sin, Ein, P, n, m, H1, H2, zt, S1, S2, zs, G1, G2)
        return f

    # ----------------------------------------------------------- backward
    def loss_grad(self, f, target, regs, P, n):
        """Centred weighted MSE (oracle.warmnet.loss) and d loss / d f."""
        df = torch.empty_like(f)
        lp = torch.empty(P, dtype=torch.float32, device=self.device)
        self._op(OP_LOSS, seg_len=n, P=P, scalar=R0, x=f, y=target, z=regs, out=df, out2=lp)
        return lp, df

    def _encode_back(self, inp, A, B, dz, rows, seg, W1, b1, W2, b2):
        dev, g, wb = self.device, self.g, self.wb
        dA2 = torch.empty((rows, H), dtype=torch.bfloat16, device=dev)
        self._op(OP_POOLBACK, rows=rows, seg_len=seg, k=H, x=dz, y=B, out=dA2)
        splits = max(1, min(64, rows // 512))
        self._gemm(A, dA2, g[W2], H, H, rows, transA=1, out_kind=2, splits=splits)
        self._op(OP_COLSUM, rows=rows, k=H, x=dA2, out=g[b2])
        dA1 = torch.empty_like(dA2)
        self._gemm(dA2, wb[W2], dA1, rows, H, H, transB=1, mask=A)
        self._gemm(inp, dA1, g[W1], KIN, H, rows, transA=1, out_kind=2, splits=splits)
        self._op(OP_COLSUM, rows=rows, k=H, x=dA1, out=g[b1])

    def backward(self, df):
        Sin, Ein, P, n, m, H1, H2, zt, S1, S2, zs, G1, G2 = self._cache
        dev = self.device
        w, wb, g = self.w, self.wb, self.g
        for t in g.values():
            t.zero_()
        rows_n, rows_m = P * n, P * m
        self._op(OP_HEADWGRAD, rows=rows_n, k=H, x=G2, y=df, out=g["w_m3"], out2=g["b_m3"])
        dB2 = torch.empty((rows_n, H), dtype=torch.bfloat16, device=dev)
        self._op(OP_HEADBACK, rows=rows_n, k=H, x=df, w=w["w_m3"], y=G2, out=dB2)
        splits = max(1, min(64, rows_n // 512))
        self._gemm(G1, dB2, g["W_m2"], H, H, rows_n, transA=1, out_kind=2, splits=splits)
        self._op(OP_COLSUM, rows=rows_n, k=H, x=dB2, out=g["b_m2"])
        dB1 = torch.empty_like(dB2)
        self._gemm(dB2, wb["W_m2"], dB1, rows_n, H, H, transB=1, mask=G1)
        self._gemm(Sin, dB1, g["W_m1x"], KIN, H, rows_n, transA=1, out_kind=2, splits=splits)
        dc = torch.empty((P, H), dtype=torch.float32, device=dev)
        self._op(OP_SEGSUM, seg_len=n, P=P, k=H, x=dB1, out=dc)
        ones = torch.ones((P, 1), dtype=torch.float32, device=dev)
        self._op(OP_OUTER, P=P, k=1, n=H, x=ones, y=dc, out=g["b_m1"])
        self._op(OP_OUTER, P=P, k=H, n=H, x=zt, y=dc, out=g["W_m1z"])
        self._op(OP_OUTER, P=P, k=H, n=H, x=zs, y=dc, out=g["W_m1s"])
        dzt = torch.empty((P, H), dtype=torch.float32, device=dev)
        self._op(OP_VECMATT, P=P, k=H, n=H, x=dc, w=w["W_m1z"], out=dzt)
        dzs = torch.empty_like(dzt)
        self._op(OP_VECMATT, P=P, k=H, n=H, x=dc, w=w["W_m1s"], out=dzs)
        self._encode_back(Ein, H1, H2, dzt, rows_m, m, "W_e1", "b_e1", "W_e2", "b_e2")
        self._encode_back(Sin, S1, S2, dzs, rows_n, n, "W_s1", "b_s1", "W_s2", "b_s2")
        return g

    def adam_step(self, lr, b1=0.9, b2=0.999, eps=1e-8, wd=0.0):
        self.step_count += 1
        t = self.step_count
        bc1, bc2 = 1 - b1 ** t, 1 - b2 ** t
        for k, wt in self.w.items():
            out2 = self.wb[k] if k in self.wb else None
            self._op(OP_ADAM, count=wt.numel(), hp=(lr, b1, b2, eps, bc1, bc2, wd),
                     x=self.g[k], y=self.m1[k], z=self.m2[k], out=wt, out2=out2)
        # padded input rows of the first layers stay exactly zero
        for k in PADDED_INPUT:
            self.w[k][DX + 2:].zero_()
            self.wb[k][DX + 2:].zero_()

    # ---------------------------------------------------------- inference
    def inputs(self, X, Y, a, b, regs, cmax, P, n, m):
        """Network inputs of P equal-size problems (X: (P*n, 16) fp32, ...)."""
        Sin = self.build_inputs(X, P, n, a, regs, cmax)
        Ein = self.build_inputs(Y, P, m, b, regs, cmax)
        qn = self.row_norms(X, P, n, cmax)
        return Sin, Ein, qn

    def predict_u0(self, R, X, Y, regs):
        """Initial log potentials u0 = f / reg for the ragged batch R (equal sizes)."""
        P = R.batch
        n, m = int(R.ns[0]), int(R.ms[0])
        if not (np.all(R.ns == n) and np.all(R.ms == m)):
            raise ValueError("the warm-start network expects problems of one size")
        regs_h = np.asarray(regs, dtype=np.float64).reshape(-1)
        # The batch's own device copy when it holds the same values: an upload
        # here would wait for the work already queued on the stream.
        R_regs_h = getattr(R, "regs_host", None)
        if (R_regs_h is not None and R.regs.device == self.device
                and R.regs.dtype == torch.float64 and np.array_equal(R_regs_h, regs_h)):
            regs_t = R.regs
        else:
            regs_t = torch.from_numpy(regs_h.copy()).to(self.device)
        cmax = R.cmax if R.cmax.dtype == torch.float32 else R.cmax.float()
        if n % 64 == 0 and m % 64 == 0:
            f = self.forward_points(X, Y, R.a, R.b, regs_t, cmax.contiguous(), P, n, m)
        else:
            Sin, Ein, qn = self.inputs(X.reshape(-1, DX).float(), Y.reshape(-1, DX).float(),
                                       R.a, R.b, regs_t, cmax, P, n, m)
            f = self.forward(Sin, Ein, qn, P, n, m)
        return (f.double().view(P, n) / regs_t[:, None]).reshape(-1)
