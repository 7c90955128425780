"""Batched docking of a ligand set against one receptor on the GPU.

One call of :meth:`DockEngine.dock` is the reference's DockTask::run
(unidock/unidock_engine/src/task/DockTask.cpp:39-123) for a whole batch:
Monte Carlo search (ud2_mc), clustering (ud2_cluster), refinement with the
penalty-slope ramp (ud2_optimize, DockTask.cpp:155-186), filtering by
energy_range / num_pose (DockTask.cpp:188-220) and final scoring with the
affinity of DockTask.cpp:224-282 (ud2_final_score, and ud2_score_decomp for the
energy decomposition). Bias terms and constraint docking are options of the
ligand set (N.LigSet). Everything but the filter runs on the device; the filter
is a sort of a few floats per ligand.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _native as N
from .topology import RAD2DEG_F32

W_CONF_INDEP = np.float32(0.05846)  # src/score/vina.h:123
V_CAP_AUTHENTIC = 1000.0            # src/constants/constants.h:19
PENALTY_SLOPE = 1e6                 # src/constants/constants.h:16
DEFAULT_PROBE_TYPES = (2, 3, 4, 5, 6, 10, 11, 15)


def heuristic_opt_steps(natom_max: int) -> int:
    """opt_steps = -1 -> (25 + natom_max) / 3 (src/screening/screening.cpp:206-208)."""
    return int((25 + natom_max) // 3)


def refine_slopes(refine_steps: int) -> list:
    """Penalty slope per refinement round, 100 * 10^(2p) capped at 1e6 (DockTask.cpp:171-175)."""
    out = []
    for p in range(refine_steps):
        s = 100.0 * 10.0 ** (2.0 * p)
        out.append(float(np.float32(min(s, PENALTY_SLOPE))))
    return out


@dataclass
class DockParams:
    nchain: int = 128            # exhaustiveness
    mc_steps: int = 20
    opt_steps: int = -1          # -1: heuristic from the largest ligand
    refine_steps: int = 5
    num_pose: int = 10
    rmsd_limit: float = 1.0
    energy_range: float = 10.0
    seed: int = 12345
    randomize: bool = True
    mode: int = N.MODE_GRID
    refine_mode: int | None = None  # None: same as mode
    search: str = "mc"           # "mc": ud2_mc; "plain" / "exact" / "learned": the learned kernel's policies
    n_cand: int = 16             # candidate mutations per MC step (exact, learned)
    top_k: int = 4               # candidates the learned policy rescores exactly
    sel_temp: float = 0.0        # learned: > 0 draws one candidate with p ~ exp(-score / sel_temp) instead
    n_start: int = 1             # > 1: randomised starts per chain, the best kept
    energy_decomp: bool = False  # per-atom inter terms of every output pose (score_decomp)


@dataclass
class LigandResult:
    name: str
    pose_ids: list = field(default_factory=list)
    coords: list = field(default_factory=list)      # [n_out] arrays (natom*3,)
    dihedrals: list = field(default_factory=list)   # [n_out] arrays (ntors,) degrees
    energies: list = field(default_factory=list)    # [n_out] [affinity, total, intra, inter, penalty, conf_indep, bias]
    decomp: list = field(default_factory=list)      # [n_out] arrays (natom, 5), with DockParams.energy_decomp


class DockEngine:
    """Receptor + box on the device, with grid maps when grid mode is used."""

    def __init__(self, receptor, box, mode: int = N.MODE_GRID, spacing: float = 0.375, probe_types=None,
                 proposal=None):
        self.receptor = receptor
        self.box = box
        self.scene = N.Scene(receptor, box)
        self.mode = mode
        self.proposal = proposal  # N.Mlp for the learned search
        if mode == N.MODE_GRID:
            self.probe_types = tuple(probe_types or DEFAULT_PROBE_TYPES)
            self.scene.build_grids(spacing, self.probe_types)

    def ensure_types(self, ligands, spacing: float = 0.375):
        """Rebuilds the maps when a ligand has an atom type without a map."""
        if self.mode != N.MODE_GRID:
            return
        need = set(int(t) for l in ligands for t in np.unique(l.types) if t != 0)
        if not need.issubset(self.probe_types):
            self.probe_types = tuple(sorted(set(self.probe_types) | need))
            self.scene.build_grids(spacing, self.probe_types)

    def dock(self, ligset: N.LigSet, prm: DockParams, poses: N.PoseBuffers | None = None, sync: bool = True,
             mc_events: list | None = None, mc_stats=None):
        """Docks every ligand of ligset; returns (list[LigandResult], PoseBuffers).

        mc_events: if given, a (start, end) pair of timing events recorded on the
        launch stream around the MC search is appended. mc_stats: optional int64
        device tensor [4] that accumulates the MC kernel's counters.
        """
        import torch
        ligs = ligset.ligands
        nchain = prm.nchain
        opt_steps = prm.opt_steps if prm.opt_steps >= 0 else heuristic_opt_steps(max(l.natom for l in ligs))
        pb = poses if poses is not None else N.PoseBuffers(ligset, nchain)
        st = N.stream_ptr()
        lib = N.lib()
        pb.init()
        mc = N.McParams(mc_steps=prm.mc_steps, opt_steps=opt_steps, randomize=int(prm.randomize),
                        seed=prm.seed & 0xFFFFFFFF, mode=prm.mode, lig_index_base=0,
                        stats_d=(mc_stats.data_ptr() if mc_stats is not None else None))
        if mc_events is not None:
            ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            ev0.record()
        if prm.search == "mc":
            N.check(lib.ud2_mc(self.scene.h, ligset.h, pb.ref, C.byref(mc), st), "ud2_mc")
        else:
            policy = {"plain": N.POLICY_PLAIN, "exact": N.POLICY_EXACT, "learned": N.POLICY_LEARNED}[prm.search]
            if policy == N.POLICY_LEARNED and self.proposal is None:
                raise N.Ud2Error("learned search needs DockEngine.proposal (a trained network)")
            lp = N.LmcParams()
            lp.mc = mc
            lp.policy, lp.n_cand, lp.top_k, lp.sel_temp = policy, prm.n_cand, prm.top_k, prm.sel_temp
            lp.n_start = prm.n_start
            net = self.proposal.h if policy == N.POLICY_LEARNED else None
            N.check(lib.ud2_lmc(self.scene.h, ligset.h, pb.ref, C.byref(lp), net, st), "ud2_lmc")
        if mc_events is not None:
            ev1.record()
            mc_events.append((ev0, ev1))
        n_lig = ligset.n_lig
        leaders = torch.zeros((n_lig, nchain), dtype=torch.int32, device="cuda")
        nlead = torch.zeros(n_lig, dtype=torch.int32, device="cuda")
        N.check(lib.ud2_cluster(ligset.h, pb.ref, float(prm.rmsd_limit), C.c_void_p(leaders.data_ptr()),
                                C.c_void_p(nlead.data_ptr()), st), "ud2_cluster")
        # pose ids of every cluster leader
        mask = torch.arange(nchain, device="cuda")[None, :] < nlead[:, None].long()
        base = torch.arange(n_lig, device="cuda", dtype=torch.int32)[:, None] * nchain
        sel = (leaders + base)[mask].to(torch.int32).contiguous()
        n_sel = int(sel.numel())
        rmode = prm.mode if prm.refine_mode is None else prm.refine_mode
        for slope in refine_slopes(prm.refine_steps):
            N.check(lib.ud2_optimize(self.scene.h, ligset.h, pb.ref, C.c_void_p(sel.data_ptr()), n_sel, rmode, 2,
                                     V_CAP_AUTHENTIC, slope, None, st), "ud2_optimize")
        # filter (DockTask.cpp:188-220) on the host: per ligand, leaders by energy
        e_all = pb.e[:n_lig * nchain].view(n_lig, nchain)
        sel_h = sel.cpu().numpy()
        e_sel = e_all.reshape(-1)[sel.long()].cpu().numpy()
        nl = nlead.cpu().numpy()
        keep_ids, owner = [], []
        pos = 0
        for l in range(n_lig):
            ids = sel_h[pos:pos + nl[l]]
            es = e_sel[pos:pos + nl[l]]
            pos += nl[l]
            if len(ids) == 0:
                continue
            order = np.argsort(es, kind="stable")
            emin = es[order[0]]
            cnt = 0
            for j in order:
                if es[j] > emin + np.float32(prm.energy_range):
                    break
                keep_ids.append(int(ids[j]))
                owner.append(l)
                cnt += 1
                if cnt >= prm.num_pose:
                    break
        keep = torch.as_tensor(np.asarray(keep_ids, np.int32), device="cuda")
        out4 = torch.zeros((max(len(keep_ids), 1), 4), device="cuda")
        if keep_ids:
            N.check(lib.ud2_final_score(self.scene.h, ligset.h, pb.ref, C.c_void_p(keep.data_ptr()), len(keep_ids),
                                        C.c_void_p(out4.data_ptr()), st), "ud2_final_score")
        stride = max(l.natom for l in ligs)
        decomp = None
        if prm.energy_decomp and keep_ids:
            decomp = torch.zeros((len(keep_ids), stride, 5), device="cuda")
            N.check(lib.ud2_score_decomp(self.scene.h, ligset.h, pb.ref, C.c_void_p(keep.data_ptr()), len(keep_ids),
                                         stride, C.c_void_p(decomp.data_ptr()), st), "ud2_score_decomp")
        if sync:
            torch.cuda.synchronize()
        results = [LigandResult(l.name) for l in ligs]
        if not sync:
            return results, pb
        n_keep = len(keep_ids)
        sc = out4[:n_keep].cpu().numpy().astype(np.float32)
        dc = decomp.cpu().numpy() if decomp is not None else None
        x_h = pb.x.cpu().numpy()
        d_deg = (pb.dihe.cpu().numpy() * RAD2DEG_F32).astype(np.float32)
        # final energies of all kept poses at once, in FP32 in the order of run_score
        # (DockTask.cpp:248-262): total = intra + inter + penalty; inter-molecular part against the
        # ligand's rank-1 intra; affinity = that / (1 + w * ntors)
        own = np.asarray(owner, np.int64)
        intra, inter, pen, bias = sc[:, 0], sc[:, 1], sc[:, 2], sc[:, 3]
        first = np.ones(n_keep, bool)
        first[1:] = own[1:] != own[:-1]
        intra_r1 = intra[np.maximum.accumulate(np.where(first, np.arange(n_keep), 0))]
        total = (intra + inter) + pen
        e_inter = total - intra_r1
        ntors = np.asarray([ligs[l].ntors for l in owner], np.float32)
        aff = e_inter / (np.float32(1) + W_CONF_INDEP * ntors)
        energies = np.stack([aff, total, intra, inter, pen, aff - e_inter, bias], axis=1).astype(np.float32)
        atom_off = ligset.arrays["atom_off"]
        tors_off = ligset.arrays["tors_off"]
        for k in range(n_keep):
            pid, l = keep_ids[k], owner[k]
            L = ligs[l]
            ch = pid - l * nchain
            r = results[l]
            xo = (int(atom_off[l]) * nchain + ch * L.natom) * 3
            do = int(tors_off[l]) * nchain + ch * L.ntors
            r.pose_ids.append(pid)
            r.coords.append(x_h[xo:xo + 3 * L.natom].copy())
            r.dihedrals.append(d_deg[do:do + L.ntors].copy())
            r.energies.append(energies[k])
            if dc is not None:
                r.decomp.append(dc[k, :L.natom].copy())
        return results, pb
