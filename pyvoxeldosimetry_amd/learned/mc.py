"""GPU Monte Carlo dose (``pvd_mc_dose`` in include/pvd_learned.h).

Produces the heterogeneous-medium dose targets of the learned correction and
the Monte-Carlo-derived point kernel that the plain convolution uses in the
learned pipeline, so that homogeneous water is consistent between the two.
Beta electrons (condensed history) and, by default, the nuclide's photons
(Lu-177 gamma lines, Y-90 bremsstrahlung) as kerma packets.
"""
from __future__ import annotations

import ctypes

import numpy as np

from .. import _native as nat
from . import spectra


class MonteCarlo:
    def __init__(self, nuclide: str = "Y90", step_mm: float = 0.5, scatter: float = 0.6, rho_min: float = 1e-3,
                 device=None, photons: bool = True, photon_every: int = 16, photon_lines=None,
                 brems_z=None, electrons: bool = True):
        """``photons``: the nuclide's gamma lines and bremsstrahlung (``spectra.PHOTON_LINES``,
        ``spectra.BREMS_Z``) as kerma packets, for every ``photon_every``-th history of a
        voxel (weight x photon_every); ``photon_lines`` / ``brems_z`` override them;
        ``electrons=False`` transports photons only."""
        import torch
        nat.require_device()
        if nuclide not in spectra.BETA_SPECTRA:
            raise ValueError(f"no beta spectrum for {nuclide}; have {list(spectra.BETA_SPECTRA)}")
        self.lib = nat.load()
        self.nuclide = nuclide
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        e_u = spectra.inverse_cdf(nuclide)
        e_r, self.r_max = spectra.energy_of_range(float(e_u[-1]))
        self.e_of_u = torch.tensor(e_u, dtype=torch.float32, device=self.device)
        self.e_of_r = torch.tensor(e_r, dtype=torch.float32, device=self.device)
        self.step_mm, self.scatter, self.rho_min = float(step_mm), float(scatter), float(rho_min)
        lines = spectra.PHOTON_LINES[nuclide] if photon_lines is None else tuple(photon_lines)
        self.photon_lines = tuple(lines) if photons else ()
        if len(self.photon_lines) > nat.PVD_MC_MAX_LINES:
            raise ValueError(f"at most {nat.PVD_MC_MAX_LINES} photon lines")
        self.brems_z = (spectra.BREMS_Z[nuclide] if brems_z is None else float(brems_z)) if photons else 0.0
        self.photon_every = int(photon_every)
        self.electrons = bool(electrons)
        mu, self.mu_lo, self.mu_hi = spectra.mu_en_table()
        self.mu_en = torch.tensor(mu, dtype=torch.float32, device=self.device)
        self.last_totals = None

    def edep(self, activity, density, voxel_mm: float, histories: int = 64, seed: int = 1, out=None):
        """Energy deposited per voxel (MeV x activity weight) as an FP32 tensor."""
        import torch
        a = activity.to(device=self.device, dtype=torch.float32).contiguous()
        d = density.to(device=self.device, dtype=torch.float32).contiguous() if density is not None else None
        if d is not None and d.shape != a.shape:
            raise ValueError("density and activity shapes differ")
        if out is None:
            out = torch.zeros(a.shape, dtype=torch.float64, device=self.device)
        elif out.dtype != torch.float64 or out.shape != a.shape or not out.is_contiguous():
            raise ValueError("out must be a contiguous float64 tensor of the activity's shape")
        totals = torch.zeros(3, dtype=torch.float64, device=self.device)
        p = nat.McParams()
        p.n[:] = tuple(a.shape)
        p.voxel_mm = float(voxel_mm)
        p.histories = int(histories)
        p.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
        p.e_of_u, p.n_u = self.e_of_u.data_ptr(), self.e_of_u.numel()
        p.e_of_r, p.n_r = self.e_of_r.data_ptr(), self.e_of_r.numel()
        p.r_max_mm = self.r_max
        p.katz_a, p.katz_b, p.katz_c = spectra.KATZ_A, spectra.KATZ_B, spectra.KATZ_C
        p.step_mm, p.scatter, p.rho_min = self.step_mm, self.scatter, self.rho_min
        p.n_lines = len(self.photon_lines)
        for i, (e, y) in enumerate(self.photon_lines):
            p.line_energy[i], p.line_yield[i] = float(e), float(y)
        p.brems_z = self.brems_z
        p.photon_every = self.photon_every
        p.electrons = int(self.electrons)
        p.mu_en, p.n_mu = self.mu_en.data_ptr(), self.mu_en.numel()
        p.mu_log_e_lo, p.mu_log_e_hi = self.mu_lo, self.mu_hi
        # emitting voxels only: one thread per (emitting voxel, history)
        active = torch.nonzero(a.reshape(-1) > 0).reshape(-1).contiguous()
        p.active, p.n_active = active.data_ptr(), active.numel()
        if active.numel() == 0:
            self.last_totals = totals
            return out
        with torch.cuda.device(self.device):
            nat.check(self.lib.pvd_mc_dose(ctypes.byref(p), ctypes.c_void_p(a.data_ptr()),
                                           ctypes.c_void_p(d.data_ptr() if d is not None else 0),
                                           ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(totals.data_ptr()),
                                           ctypes.c_void_p(nat.current_stream_handle(self.device))), "pvd_mc_dose")
        self.last_totals = totals
        return out

    def dose(self, activity, density, voxel_mm: float, histories: int = 64, seed: int = 1):
        """Dose up to a constant: deposited energy / (rho * voxel volume), rho in g/cc (FP32)."""
        import torch
        e = self.edep(activity, density, voxel_mm, histories, seed)
        rho = density.to(self.device, torch.float64) if density is not None else torch.ones_like(e)
        d = torch.where(rho > 0, e / torch.clamp(rho, min=self.rho_min), torch.zeros_like(e))
        return d.float()

    def point_kernel(self, voxel_mm: float, size: int, histories: int = 1 << 22, seed: int = 7):
        """Energy-deposition kernel of a unit source in the centre voxel of a water cube,
        symmetrised over the 48 symmetries of the cube (FP64 NumPy, sums to ~1 minus escape)."""
        import torch
        if size % 2 == 0:
            raise ValueError("kernel size must be odd")
        a = torch.zeros((size, size, size), dtype=torch.float32, device=self.device)
        c = size // 2
        a[c, c, c] = 1.0
        k = self.edep(a, None, voxel_mm, histories, seed).cpu().numpy()
        sym = np.zeros_like(k)
        for perm in ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)):
            kp = np.transpose(k, perm)
            for fx in (False, True):
                for fy in (False, True):
                    for fz in (False, True):
                        t = kp
                        if fx:
                            t = t[::-1, :, :]
                        if fy:
                            t = t[:, ::-1, :]
                        if fz:
                            t = t[:, :, ::-1]
                        sym += t
        return sym / 48.0
