"""Learned orientation proposal: a CNN (bsx_cnn_* kernels) trained on
projections that the library itself extracts from the map, proposing the
top-k coarse orientation classes of each particle; the library views of
those classes are scored by the FP32 FRC, refined locally, and particles
whose top proposal is not confident fall back to the exhaustive search.

No counterpart in the reference: its exhaustive search
(part_determine_orientation, src/mg/mg_orient.cpp:545-798) stays the
fallback and the yardstick of angular error.
"""
from __future__ import annotations

import ctypes as C
import math
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import synth
from .abi import call, lib
from .recon import _ptr, _stream, central_sections, fft
from .search import Search, SearchResult, view_matrices, views_c1


# ------------------------------------------------------------ classes
class OrientationClasses:
    """Coarse SO(3) classes: the C1 view grid at `step` degrees times in-plane
    angles at `step` degrees. Class (d, q) is the rotation of view d with its
    in-plane angle advanced by q * step."""

    def __init__(self, step_deg: float):
        self.step = math.radians(step_deg)
        base = views_c1(step_deg)
        self.ndir = len(base)
        self.nang = int(round(2 * math.pi / self.step))
        self.dirs = base[:, :3].copy()
        self.base = base
        views = np.repeat(base, self.nang, axis=0)
        views[:, 3] += np.tile(np.arange(self.nang) * self.step, self.ndir)
        self.mats = view_matrices(views)                       # (C, 3, 3)
        self.base_mats = view_matrices(base)
        self.n = len(self.mats)

    def assign(self, mats: np.ndarray, chunk: int = 65536) -> np.ndarray:
        """Nearest class (geodesic) of each rotation: the best of 4 nearest
        directions x 3 in-plane bins around the in-plane angle estimate."""
        mats = np.asarray(mats, dtype=np.float64).reshape(-1, 3, 3)
        out = np.empty(len(mats), dtype=np.int64)
        for s in range(0, len(mats), chunk):
            m = mats[s:s + chunk]
            v = m[:, :, 2]                                     # view direction R[:, 2]
            d = v @ self.dirs.T                                # (n, D)
            top = np.argpartition(-d, min(3, self.ndir - 1), axis=1)[:, :4]
            best = np.full(len(m), -np.inf)
            cls = np.zeros(len(m), dtype=np.int64)
            for j in range(top.shape[1]):
                dj = top[:, j]
                rel = np.einsum("nji,njk->nik", self.base_mats[dj], m)
                q0 = np.rint(np.arctan2(rel[:, 1, 0], rel[:, 0, 0]) / self.step).astype(np.int64)
                for dq in (-1, 0, 1):
                    q = (q0 + dq) % self.nang
                    c = dj * self.nang + q
                    tr = np.einsum("nij,nij->n", self.mats[c], m)
                    better = tr > best
                    best = np.where(better, tr, best)
                    cls = np.where(better, c, cls)
            out[s:s + chunk] = cls
        return out

    def members(self, lib_mats: np.ndarray) -> np.ndarray:
        """(C, maxm) table of library view indices per class, -1 padded."""
        a = self.assign(lib_mats)
        order = np.argsort(a, kind="stable")
        counts = np.bincount(a, minlength=self.n)
        maxm = int(counts.max())
        tab = np.full((self.n, maxm), -1, dtype=np.int32)
        starts = np.concatenate([[0], np.cumsum(counts)[:-1]])
        pos = np.arange(len(a)) - np.repeat(starts, counts)
        tab[a[order], pos] = order
        return tab


# ------------------------------------------------------------ network
class ProposalNet:
    """The 5-layer CNN of bsx_cnn_* (input size x size, four conv layers, FC)."""

    def __init__(self, size: int, nclasses: int, channels=(32, 64, 128, 128), max_batch: int = 512,
                 seed: int = 1, device="cuda"):
        self.size, self.nclasses, self.channels = size, nclasses, tuple(channels)
        self.max_batch = max_batch
        self.device = torch.device(device)
        ch = (C.c_int32 * 4)(*channels)
        self.h = C.c_void_p()
        call("bsx_cnn_create", C.byref(self.h), size, ch, nclasses, max_batch, seed)
        n, f = C.c_int64(), C.c_int64()
        call("bsx_cnn_nparams", self.h, C.byref(n), C.byref(f))
        self.nparams, self.features = n.value, f.value

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value:
            try:
                lib().bsx_cnn_destroy(h)
            except Exception:
                pass
            self.h = None

    def params(self) -> torch.Tensor:
        out = torch.empty(self.nparams, dtype=torch.float32, device=self.device)
        call("bsx_cnn_params", self.h, _ptr(out), None, _stream())
        return out

    def set_params(self, p: torch.Tensor) -> None:
        p = p.to(device=self.device, dtype=torch.float32).contiguous()
        assert p.numel() == self.nparams
        call("bsx_cnn_params", self.h, None, _ptr(p), _stream())

    def forward(self, x: torch.Tensor) -> torch.Tensor:
        x = x.contiguous()
        out = torch.empty((len(x), self.nclasses), dtype=torch.float32, device=self.device)
        call("bsx_cnn_forward", self.h, _ptr(x), len(x), _ptr(out), _stream())
        return out

    def gradients(self, x: torch.Tensor, labels: torch.Tensor):
        x = x.contiguous()
        labels = labels.to(device=self.device, dtype=torch.int32).contiguous()
        g = torch.empty(self.nparams, dtype=torch.float32, device=self.device)
        loss = torch.empty(len(x), dtype=torch.float32, device=self.device)
        call("bsx_cnn_gradients", self.h, _ptr(x), _ptr(labels), len(x), _ptr(g), _ptr(loss),
             _stream())
        return g, loss

    def train_step(self, x: torch.Tensor, labels: torch.Tensor, lr: float,
                   weight_decay: float = 0.0) -> torch.Tensor:
        loss = torch.empty(len(x), dtype=torch.float32, device=self.device)
        call("bsx_cnn_train_step", self.h, _ptr(x), _ptr(labels), len(x), float(lr),
             float(weight_decay), _ptr(loss), _stream())
        return loss

    def topk(self, x: torch.Tensor, k: int):
        x = x.contiguous()
        cls = torch.empty((len(x), k), dtype=torch.int32, device=self.device)
        prob = torch.empty((len(x), k), dtype=torch.float32, device=self.device)
        call("bsx_cnn_infer_topk", self.h, _ptr(x) if len(x) else None, len(x), int(k),
             _ptr(cls) if len(x) else None, _ptr(prob) if len(x) else None, _stream())
        return cls, prob

    def fit(self, x: torch.Tensor, labels: torch.Tensor, epochs: int = 4, batch: int = 256,
            lr: float = 2e-3, weight_decay: float = 0.0, seed: int = 0, log=None):
        """Adam over shuffled minibatches, cosine-decayed learning rate."""
        labels = labels.to(device=self.device, dtype=torch.int32).contiguous()
        n = len(x)
        g = torch.Generator(device="cpu")
        g.manual_seed(seed)
        steps = epochs * (n // batch)
        it = 0
        hist = []
        for ep in range(epochs):
            perm = torch.randperm(n, generator=g).to(self.device)
            tot = torch.zeros((), device=self.device)
            for s in range(0, n - batch + 1, batch):
                idx = perm[s:s + batch]
                xb = x[idx].contiguous()
                yb = labels[idx].contiguous()
                cur = lr * 0.5 * (1 + math.cos(math.pi * it / max(1, steps)))
                tot += self.train_step(xb, yb, cur, weight_decay).sum()
                it += 1
            hist.append(tot.item() / max(1, (n // batch) * batch))
            if log:
                log(f"epoch {ep + 1}/{epochs}: mean loss {hist[-1]:.4f}")
        return hist


class EmbeddingNet(ProposalNet):
    """The same network read as an embedding of `dim` values (bsx_cnn_embed):
    trained contrastively on (particle, clean projection) pairs of many maps
    (bsx_cnn_train_pairs), it proposes orientations for a map it has not seen
    by comparing each particle with that map's own class projections."""

    def __init__(self, size: int, dim: int = 64, channels=(32, 64, 128, 128),
                 max_batch: int = 512, seed: int = 1, device="cuda"):
        super().__init__(size, dim, channels, max_batch, seed, device)
        self.dim = dim

    def embed(self, x: torch.Tensor) -> torch.Tensor:
        x = x.contiguous()
        out = torch.empty((len(x), self.dim), dtype=torch.float32, device=self.device)
        if len(x):
            call("bsx_cnn_embed", self.h, _ptr(x), len(x), _ptr(out), _stream())
        return out

    def train_pairs(self, xp: torch.Tensor, xr: torch.Tensor, tau: float, lr: float,
                    weight_decay: float = 0.0) -> torch.Tensor:
        x = torch.cat([xp, xr]).contiguous()
        loss = torch.empty(len(xp), dtype=torch.float32, device=self.device)
        call("bsx_cnn_train_pairs", self.h, _ptr(x), len(xp), float(tau), float(lr),
             float(weight_decay), _ptr(loss), _stream())
        return loss

    def pair_gradients(self, xp: torch.Tensor, xr: torch.Tensor, tau: float):
        """(flat gradient, per-pair losses) of the contrastive loss, no update."""
        x = torch.cat([xp, xr]).contiguous()
        g = torch.empty(self.nparams, dtype=torch.float32, device=self.device)
        loss = torch.empty(len(xp), dtype=torch.float32, device=self.device)
        call("bsx_cnn_pair_gradients", self.h, _ptr(x), len(xp), float(tau), _ptr(g), _ptr(loss),
             _stream())
        return g, loss

    def match_topk(self, p: torch.Tensor, r: torch.Tensor, k: int, tau: float):
        """The k best references (rows of r) of each particle embedding (rows of p)."""
        p, r = p.contiguous(), r.contiguous()
        cls = torch.empty((len(p), k), dtype=torch.int32, device=self.device)
        prob = torch.empty((len(p), k), dtype=torch.float32, device=self.device)
        if len(p):
            call("bsx_cnn_match_topk", self.h, _ptr(p), len(p), _ptr(r), len(r), int(k), float(tau),
                 _ptr(cls), _ptr(prob), _stream())
        return cls, prob

    def fit_pairs(self, xp: torch.Tensor, xr: torch.Tensor, epochs: int = 4, batch: int = 128,
                  lr: float = 2e-3, tau: float = 0.1, weight_decay: float = 0.0, seed: int = 0,
                  log=None):
        """Adam over shuffled pair minibatches, cosine-decayed learning rate."""
        n = len(xp)
        g = torch.Generator(device="cpu")
        g.manual_seed(seed)
        steps = epochs * (n // batch)
        it = 0
        hist = []
        for ep in range(epochs):
            perm = torch.randperm(n, generator=g).to(self.device)
            tot = torch.zeros((), device=self.device)
            for s in range(0, n - batch + 1, batch):
                idx = perm[s:s + batch]
                cur = lr * 0.5 * (1 + math.cos(math.pi * it / max(1, steps)))
                tot += self.train_pairs(xp[idx], xr[idx], tau, cur, weight_decay).sum()
                it += 1
            hist.append(tot.item() / max(1, (n // batch) * batch))
            if log:
                log(f"epoch {ep + 1}/{epochs}: mean pair loss {hist[-1]:.4f}")
        return hist


# ------------------------------------------------------------ training data
def extracted_projections(ft: torch.Tensor, box: int, mats: np.ndarray, rng: np.random.Generator,
                          *, pix=1.5, volt=300e3, cs=2.7e7, amp_contrast=0.07,
                          defocus=(1e4, 3e4), snr=0.1, kernel_width=4, chunk=2048, ctf=True):
    """Particle images made from central sections that the library extracts
    from the map transform (bsx_extract, FSI kernel), with a per-particle CTF
    (sin chi, unless ctf is False) and white noise of the given SNR (none when
    snr is None). Returns (images (n, box, box) float32 on the device,
    PARTICLE_DTYPE records)."""
    n = len(mats)
    parts = np.zeros(n, dtype=synth.PARTICLE_DTYPE)
    parts["ori"] = box // 2
    parts["hdr_ori"] = box // 2
    parts["pixel_size"] = pix
    parts["mag"] = 1.0
    parts["volt"] = volt
    parts["cs"] = cs
    parts["amp_shift"] = math.asin(amp_contrast)
    parts["def_avg"] = rng.uniform(*defocus, n)
    parts["sel"] = 1.0
    dev = ft.device
    imgs = torch.empty((n, box, box), dtype=torch.float32, device=dev)
    fh = torch.fft.fftfreq(box, d=1.0 / box, device=dev).to(torch.float64)
    hh, kk = fh[None, None, :], fh[None, :, None]
    gen = torch.Generator(device=dev)
    gen.manual_seed(int(rng.integers(1 << 62)))
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        sec = central_sections(ft, mats[s:e].reshape(-1, 9), pix, 2 * pix, kernel_width, 2, True)
        if ctf:
            dfa = torch.as_tensor(parts["def_avg"][s:e], device=dev, dtype=torch.float64)[:, None, None]
            zero = torch.zeros_like(dfa)
            c = synth.ctf_sin_chi(hh, kk, box, pix, volt, cs, math.asin(amp_contrast), dfa, zero, zero)
            sec = sec * c.to(torch.float32)
        sec = sec.contiguous()
        fft(sec, 2, +1, 0)
        img = sec.real.contiguous()
        if snr is None:
            imgs[s:e] = img
            continue
        sd = img.reshape(e - s, -1).std(dim=1)
        noise = torch.randn(img.shape, generator=gen, device=dev, dtype=torch.float32)
        imgs[s:e] = img + noise * (sd / math.sqrt(snr))[:, None, None]
    return imgs, parts


def training_set(ft: torch.Tensor, box: int, n: int, classes: OrientationClasses, s: Search,
                 size: int, rng: np.random.Generator, chunk: int = 8192, **kw):
    """(inputs (n, size, size), labels (n,), rotations (n, 3, 3)) of n random
    orientations; inputs go through the same staging as real particles."""
    views = synth.uniform_views(n, rng)
    mats = view_matrices(views)
    labels = classes.assign(mats)
    xs = torch.empty((n, size, size), dtype=torch.float32, device=ft.device)
    for a in range(0, n, chunk):
        b = min(n, a + chunk)
        imgs, parts = extracted_projections(ft, box, mats[a:b], rng, **kw)
        s.stage(imgs, parts)
        xs[a:b] = s.cnn_input(size)
    return xs, torch.as_tensor(labels, dtype=torch.int32, device=ft.device), mats


def pair_set(ft: torch.Tensor, box: int, n: int, s_part: Search, s_ref: Search, size: int,
             rng: np.random.Generator, chunk: int = 8192, **kw):
    """n (particle, reference) network-input pairs of one map at random
    orientations: the particle with CTF and noise staged through s_part (a
    handle with the CTF flag: its input is CTF-flipped), the reference the
    clean projection at the same orientation staged through s_ref (no CTF)."""
    views = synth.uniform_views(n, rng)
    mats = view_matrices(views)
    xp = torch.empty((n, size, size), dtype=torch.float32, device=ft.device)
    xr = torch.empty_like(xp)
    for a in range(0, n, chunk):
        b = min(n, a + chunk)
        imgs, parts = extracted_projections(ft, box, mats[a:b], rng, **kw)
        s_part.stage(imgs, parts)
        xp[a:b] = s_part.cnn_input(size)
        kw_ref = dict(kw)
        kw_ref.update(snr=None, ctf=False)
        imgs, parts = extracted_projections(ft, box, mats[a:b], rng, **kw_ref)
        s_ref.stage(imgs, parts)
        xr[a:b] = s_ref.cnn_input(size)
    return xp, xr, mats


def class_references(net: "EmbeddingNet", ft: torch.Tensor, box: int, classes: OrientationClasses,
                     s_ref: Search, size: int, chunk: int = 8192, **kw) -> torch.Tensor:
    """Embeddings of a map's clean projections at every class rotation
    (classes.n x dim): what a particle of that map is compared with."""
    rng = np.random.default_rng(0)
    out = []
    kw = dict(kw)
    kw.update(snr=None, ctf=False)
    for a in range(0, classes.n, chunk):
        b = min(classes.n, a + chunk)
        imgs, parts = extracted_projections(ft, box, classes.mats[a:b], rng, **kw)
        s_ref.stage(imgs, parts)
        out.append(net.embed(s_ref.cnn_input(size)))
    return torch.cat(out)


# ------------------------------------------------------------ pipeline
@dataclass
class LearnedStats:
    fallback: int = 0
    candidates: int = 0
    times: dict = field(default_factory=dict)


def learned_search(s: Search, net: ProposalNet, *, k: int = 16, min_prob: float = 0.0,
                   steps_deg=(2.0, 1.0), timer=None, refs: torch.Tensor | None = None,
                   tau: float = 0.1):
    """Learned pipeline on the staged particles: top-k classes from the
    network -> BF16 GEMM against the member views of those classes only
    (routed, grouped) -> FP32 FRC of the best candidates -> local refinement.
    The classes come from the classifier's softmax, or, with `refs` (the
    class embeddings of the map, class_references), from the embedding
    network's cosine matches. Particles whose top probability is below
    min_prob also get the exhaustive candidates. The classes' member table
    must be registered with s.set_classes."""
    st = LearnedStats()
    t = timer or (lambda name: None)
    x = s.cnn_input(net.size)
    t("cnn_input")
    if refs is None:
        cls, prob = net.topk(x, k)
    else:
        cls, prob = net.match_topk(net.embed(x), refs, k, tau)
    t("cnn")
    cand, _ = s.proposed(cls)
    t("proposed_gemm")
    if min_prob > 0:
        low = torch.nonzero(prob[:, 0] < min_prob).flatten()
        st.fallback = int(low.numel())
        if st.fallback:
            ec, _ = s.exhaustive_rows(low.to(torch.int32))
            extra = torch.full_like(cand, -1)
            extra[low] = ec
            cand = torch.cat([cand, extra], dim=1).contiguous()
        t("fallback")
    st.candidates = cand.shape[1]
    _, best, bf = s.rescore(cand)
    t("rescore")
    mats = s.view_mats(best)
    if steps_deg:
        mats, frc = s.refine(mats, steps_deg)
    else:
        frc = bf
    t("refine")
    return SearchResult(mats=mats, frc=frc, grid_view=best, grid_frc=bf), st, cls, prob
