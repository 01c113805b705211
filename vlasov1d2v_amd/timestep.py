"""Shu-Osher SSP-RK3 stepping on the GPU: drop-in for the reference's
``Stepper`` (main/timestep.py:42-242, nonlinear path).

``nonlinear_ssp_rk`` runs one fused step through ``vl_rk3_step`` (7 kernel
launches, no host synchronisation).  ``nonlinear_ssp_rk_unfused`` composes
the per-operator entry points in the reference's order and is kept for
cross-checking.  With a fixed ``dt`` the loop never synchronises and can be
captured in a HIP graph.  With ``dt=None`` it follows the reference's CFL rule
(main/timestep.py:184,231-242): ``main_loop`` evaluates the rule on the
device after every step (``vl_rk3_step_adaptive``), so adaptive loops are
graph-captured too and read the clock back once per graph replay; a single
``nonlinear_ssp_rk`` call applies it on the host (one read of max|acc E|).
"""
import gc
import time as _time

import numpy as np

from . import _lib
from .grid import STORAGE_CODES, _check_state, storage_of

# Shu-Osher rows [a_fn, a_prev, a_dtL] (main/timestep.py:35-39)
NONLINEAR_SSP_RK = {
    2: [[1 / 2, 1 / 2, 1 / 2]],
    3: [[3 / 4, 1 / 4, 1 / 4],
        [1 / 3, 2 / 3, 2 / 3]],
}

# Linear SSP-RK coefficients by order (main/timestep.py:11-21)
SSP_RK_LINEAR = {
    1: [1],
    2: [1 / 2, 1 / 2],
    3: [1 / 3, 1 / 2, 1 / 6],
    4: [3 / 8, 1 / 3, 1 / 4, 1 / 24],
    5: [11 / 30, 3 / 8, 1 / 6, 1 / 12, 1 / 120],
    6: [53 / 144, 11 / 30, 3 / 16, 1 / 18, 1 / 48, 1 / 720],
    7: [103 / 280, 53 / 144, 11 / 60, 3 / 48, 1 / 72, 1 / 240, 1 / 5040],
    8: [2119 / 5760, 103 / 280, 53 / 288, 11 / 180, 1 / 64, 1 / 360, 1 / 1440, 1 / 40320],
}

# Cockburn-Shu Courant numbers [time_order][space_order - 1] (main/timestep.py:24-33)
COURANT_NUMBERS = {
    1: [1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0],
    2: [1.0, 0.333],
    3: [1.256, 0.409, 0.209, 0.130, 0.089, 0.066, 0.051, 0.040, 0.033],
    4: [1.392, 0.464, 0.235, 0.145, 0.100, 0.073, 0.056, 0.045, 0.037],
    5: [1.608, 0.534, 0.271, 0.167, 0.115, 0.085, 0.065, 0.052, 0.042],
    6: [1.776, 0.592, 0.300, 0.185, 0.127, 0.093, 0.072, 0.057, 0.047],
    7: [1.977, 0.659, 0.333, 0.206, 0.142, 0.104, 0.080, 0.064, 0.052],
    8: [2.156, 0.718, 0.364, 0.225, 0.154, 0.114, 0.087, 0.070, 0.057],
}


def get_max_speeds(grids, elliptic, acc):
    """|max speeds| per axis (main/timestep.py:235-242); one device->host read."""
    max_acc_e = float(elliptic.max_abs_acc_e(grids.x, acc).item())
    b = elliptic.magnetic_field
    return np.absolute(np.array([grids.u.arr_max,
                                 max_acc_e + abs(acc * b * grids.v.arr_max),
                                 acc * abs(b * grids.u.arr_max)]))


class Stepper:
    def __init__(self, time_order, space_order, write_time, final_time, linear=False, dt=None):
        self.linear = bool(linear)
        if self.linear:
            if time_order not in SSP_RK_LINEAR:
                raise ValueError(f"no linear SSP-RK coefficients of order {time_order}")
        elif time_order != 3:
            raise NotImplementedError("the fused GPU step implements the SSP-RK3 row only")
        self.time_order = int(time_order)
        self.space_order = int(space_order)
        self.coefficients = np.array(SSP_RK_LINEAR[self.time_order] if self.linear
                                     else NONLINEAR_SSP_RK[self.time_order])
        # The reference's table stops at 9 nodes (IndexError there for 10);
        # here a missing entry only matters for the adaptive (CFL) time step.
        row = COURANT_NUMBERS[self.time_order]
        self.courant = row[self.space_order - 1] if self.space_order - 1 < len(row) else None
        if self.courant is None and dt is None:
            raise ValueError(f"no Courant number for {self.space_order} nodes: pass a fixed dt")
        self.time = 0.0
        self.fixed_dt = dt
        self.dt = dt
        self.steps_counter = 0
        self.write_counter = 1
        self.write_time = write_time
        self.final_time = final_time
        self.time_array = np.array([self.time])
        self.field_energy = np.array([])
        self._stages = None
        self._rho_token = None
        self.elapsed = None
        self._res = None          # resident element-major state: dict or None

    # eager steps between two reads of the workspace status in a time loop
    CHECK_EVERY = 64

    # -- helpers -----------------------------------------------------------
    def check_finite(self, distribution, grids, dg_flux, elliptic):
        """Raise VlasovError if a step of this stepper met a non-finite field
        (the workspace's VL_STATUS_NONFINITE bit; one device synchronisation).
        The time loops call it after every graph replay, every CHECK_EVERY
        eager steps and at their end; single nonlinear_ssp_rk calls do not."""
        g, arr, _, _, _ = self._step_state(distribution, grids, dg_flux, elliptic)
        g.check_finite(arr.device, f"SSP-RK3 step (steps done: {self.steps_counter})")

    def _stage_buffers(self, arr):
        import torch
        if (self._stages is None or self._stages[0].shape != arr.shape
                or self._stages[0].device != arr.device):
            self._stages = (torch.empty_like(arr), torch.empty_like(arr))
        return self._stages

    def adapt_time_step(self, max_speeds, dx, du, dv):     # main/timestep.py:231-232
        self.dt = self.courant / ((max_speeds[0] / dx) + (max_speeds[1] / du) + (max_speeds[2] / dv))

    def _acc(self, dg_flux, refs):
        if refs is not None and hasattr(refs, "electron_acceleration_multiplier"):
            return refs.electron_acceleration_multiplier
        return dg_flux.acceleration_coefficient

    def _adapt(self, grids, elliptic, dg_flux, refs):
        self.adapt_time_step(get_max_speeds(grids, elliptic, self._acc(dg_flux, refs)),
                             grids.x.dx, grids.u.dx, grids.v.dx)

    # -- resident element-major stepping ------------------------------------
    def resident_begin(self, distribution, grids, dg_flux, elliptic, layout=None):
        """Copy distribution.arr into an element-wise layout (each element's
        n^3 nodes contiguous; by default the one the library reports fastest
        for this storage and n) and step there until resident_end(); the
        reference-layout distribution.arr is stale in between."""
        import torch
        arr = distribution.arr
        _check_state(arr, grids.shape, "resident_begin", allow_reduced=True)
        g = dg_flux._geom(grids, elliptic, STORAGE_CODES[storage_of(arr)])
        if layout is None:
            layout = g.resident_layout()
        n_em = g.state_elems_layout(layout)
        r = self._res
        if r is None or r["bufs"][0].numel() != n_em or r["bufs"][0].dtype != arr.dtype \
                or r["bufs"][0].device != arr.device:
            bufs = tuple(torch.zeros(n_em, dtype=arr.dtype, device=arr.device) for _ in range(3))
        else:
            bufs = r["bufs"]
        g.convert_layout(arr, _lib.VL_LAYOUT_REFERENCE, bufs[0], layout)
        self._res = {"bufs": bufs, "dist": distribution, "geom": g, "token": None, "active": True,
                     "layout": layout}

    def resident_end(self, distribution):
        """Write the resident state back into distribution.arr (reference layout)."""
        r = self._res
        if r is None or not r["active"]:
            return
        r["geom"].convert_layout(r["bufs"][0], r["layout"], distribution.arr, _lib.VL_LAYOUT_REFERENCE)
        r["active"] = False
        self._rho_token = None

    # -- one step ------------------------------------------------------------
    def nonlinear_ssp_rk(self, distribution, basis, elliptic, grids, dg_flux, refs=None, energy_out=None):
        """One SSP-RK3 step (main/timestep.py:141-187), fused on the GPU.

        On return ``elliptic.electric_field``/``potential`` hold the field of
        the third stage's input (as in the reference) and, when given,
        ``energy_out`` (a 0-d/1-element GPU tensor) its field energy.
        """
        if self.dt is None:
            self._adapt(grids, elliptic, dg_flux, refs)
        self._fused_step(distribution, elliptic, grids, dg_flux, energy_out)
        if self.fixed_dt is None:
            self._adapt(grids, elliptic, dg_flux, refs)

    def _step_state(self, distribution, grids, dg_flux, elliptic):
        """(geometry, state array the step works on, its stage buffers, layout
        flags): the resident buffers while resident_begin is active for this
        distribution, else distribution.arr in the reference layout."""
        arr = distribution.arr
        _check_state(arr, grids.shape, "nonlinear_ssp_rk", allow_reduced=True)
        g = dg_flux._geom(grids, elliptic, STORAGE_CODES[storage_of(arr)])
        res = self._res if (self._res is not None and self._res["active"]
                            and self._res["dist"] is distribution) else None
        if res is not None:
            arr, f1, f2 = res["bufs"]
            lay = (_lib.VL_STEP_ROW_SPLIT if res["layout"] == _lib.VL_LAYOUT_ROW_SPLIT
                   else _lib.VL_STEP_ELEMENT_MAJOR)
        else:
            f1, f2 = self._stage_buffers(arr)
            lay = 0
        return g, arr, f1, f2, lay

    @staticmethod
    def _token(arr, ws):
        """What must be unchanged since a step for its moment partials, left in
        the workspace, to describe ``arr``: the array, its torch version, the
        workspace and the count of library calls that write partials or states."""
        return (arr.data_ptr(), arr._version, ws.data_ptr(), _lib.generation())

    def _fused_step(self, distribution, elliptic, grids, dg_flux, energy_out=None, clock_out=None,
                    adaptive=False):
        """vl_rk3_step with self.dt, or (adaptive) vl_rk3_step_adaptive with
        the workspace clock; clock_out: 2-element GPU tensor or None."""
        import torch
        g, arr, f1, f2, lay = self._step_state(distribution, grids, dg_flux, elliptic)
        dev = arr.device
        if elliptic.electric_field is None or elliptic.electric_field.device != dev:
            elliptic.electric_field = torch.zeros((grids.x.res_ghosts, grids.x.order), dtype=torch.float64,
                                                  device=dev)
        if elliptic.potential is None or elliptic.potential.device != dev:
            elliptic.potential = torch.zeros((grids.x.res, grids.x.order), dtype=torch.float64, device=dev)
        ws = g.workspace(dev)
        flags = lay | (_lib.VL_STEP_RHO_VALID if self._token(arr, ws) == self._rho_token else 0)
        if adaptive:
            _lib.check(_lib.load().vl_rk3_step_adaptive(
                g.handle, _lib.ptr(arr), _lib.ptr(f1), _lib.ptr(f2), _lib.ptr(elliptic.electric_field),
                _lib.ptr(elliptic.potential), _lib.ptr(energy_out), _lib.ptr(clock_out), _lib.ptr(ws), flags,
                _lib.stream_ptr()), "vl_rk3_step_adaptive")
        else:
            _lib.check(_lib.load().vl_rk3_step(
                g.handle, _lib.ptr(arr), _lib.ptr(f1), _lib.ptr(f2), _lib.ptr(elliptic.electric_field),
                _lib.ptr(elliptic.potential), float(self.dt), _lib.ptr(energy_out), _lib.ptr(ws), flags,
                _lib.stream_ptr()), "vl_rk3_step")
        self._rho_token = self._token(arr, ws)

    def ssp_rk_update(self, distribution, basis, grids, elliptic, dg_flux, refs=None):
        """The linear SSP-RK update (main/timestep.py:189-229) on the GPU, in
        the reference's order, with the central-flux Poisson solve
        (``elliptic.poisson``; build_central_flux_operator and invert first).
        FP64 state in the reference layout."""
        arr = distribution.arr
        _check_state(arr, grids.shape, "ssp_rk_update")
        if self.dt is None:
            self._adapt(grids, elliptic, dg_flux, refs)
        c = self.coefficients
        geo = distribution._geom()
        lib = _lib.load()

        def field_of(copy):
            _lib.check(lib.vl_swap_ghosts(geo.handle, _lib.ptr(copy), _lib.stream_ptr()), "vl_swap_ghosts")
            density = distribution.moment_zero_of_arr(copy)
            elliptic.poisson(_lib.density_deviation(geo, density), grids.x, basis.b1)

        import torch
        distribution.swap_ghost_cells_of_array()
        copy = torch.empty_like(arr)
        _lib.axpby(geo, copy, arr, 0.0, 1.0, interior=False)    # copy_for_explicit_advance (main/grid.py:198-199)
        _lib.axpby(geo, arr, None, c[0], 0.0)                    # arr[ng] = c0 arr[ng]
        for i in range(1, self.time_order):
            field_of(copy)
            _lib.axpby(geo, copy, dg_flux.semi_discrete_rhs(copy, elliptic, basis, grids), 1.0, self.dt)
            _lib.axpby(geo, arr, copy, 1.0, c[i])
        field_of(copy)
        _lib.axpby(geo, arr, dg_flux.semi_discrete_rhs(copy, elliptic, basis, grids), 1.0, c[-1] * self.dt)
        self._rho_token = None
        if self.fixed_dt is None:
            self._adapt(grids, elliptic, dg_flux, refs)

    def nonlinear_ssp_rk_unfused(self, distribution, basis, elliptic, grids, dg_flux, refs=None):
        """The same step composed from the per-operator GPU entry points in the
        reference's order (main/timestep.py:141-187)."""
        import torch
        arr = distribution.arr
        geo = distribution._geom()
        lib = _lib.load()
        distribution.swap_ghost_cells_of_array()
        # every stage starts as a copy of f^n: its u/v ghost planes are the
        # frozen ones of f^n; its x ghosts are swapped before use
        stages = torch.empty((self.time_order,) + tuple(arr.shape), dtype=arr.dtype, device=arr.device)
        for i in range(self.time_order):
            _lib.axpby(geo, stages[i], arr, 0.0, 1.0, interior=False)
        density = distribution.moment_zero()
        elliptic.poisson2(_lib.density_deviation(geo, density), grids.x, basis.b1)
        _lib.axpby(geo, stages[0], dg_flux.semi_discrete_rhs(arr, elliptic, basis, grids), 1.0, self.dt)
        for i in range(1, self.time_order):
            _lib.check(lib.vl_swap_ghosts(geo.handle, _lib.ptr(stages[i - 1]), _lib.stream_ptr()), "vl_swap_ghosts")
            density = distribution.moment_zero_of_arr(stages[i - 1])
            elliptic.poisson2(_lib.density_deviation(geo, density), grids.x, basis.b1)
            df_dt = dg_flux.semi_discrete_rhs(stages[i - 1], elliptic, basis, grids)
            c = self.coefficients[i - 1]
            _lib.axpby(geo, stages[i], arr, 0.0, c[0])
            _lib.axpby(geo, stages[i], stages[i - 1], 1.0, c[1])
            _lib.axpby(geo, stages[i], df_dt, 1.0, c[2] * self.dt)
        _lib.axpby(geo, arr, stages[-1], 0.0, 1.0)
        self._rho_token = None

    # -- time loop -------------------------------------------------------------
    def main_loop(self, distribution, basis, elliptic, grids, dg_flux, refs=None, save_file=None,
                  steps=None, graph_steps=0, device_dt=True):
        """Step until ``final_time`` (or exactly ``steps`` steps) and record the
        field-energy series as the reference does (main/timestep.py:81-139).

        The energies (and, adaptive, the times) stay on the GPU until the loop
        ends, and with ``graph_steps`` > 0 the loop replays a HIP graph of
        that many steps (after one eager step; the remainder runs eagerly).
        With ``dt=None`` the CFL rule runs on the device (``device_dt``);
        ``device_dt=False`` keeps the host rule of ``nonlinear_ssp_rk``.
        """
        import torch
        t0 = _time.time()
        dev = distribution.arr.device
        e0 = elliptic.electric_energy(grids.x)
        if self.linear:
            self._main_loop_linear(distribution, basis, elliptic, grids, dg_flux, refs, steps, e0)
        elif self.fixed_dt is None and device_dt and self.courant is not None \
                and self._acc(dg_flux, refs) == dg_flux.acceleration_coefficient:
            self._main_loop_device_dt(distribution, basis, elliptic, grids, dg_flux, refs, steps, graph_steps, e0)
        elif self.fixed_dt is None:
            self._adapt(grids, elliptic, dg_flux, refs)
            self.field_energy = np.append(self.field_energy, float(e0.item()))
            self.resident_begin(distribution, grids, dg_flux, elliptic)
            done = 0
            while (self.time < self.final_time) if steps is None else (done < steps):
                done += 1
                energy = torch.empty((), dtype=torch.float64, device=dev)
                self.nonlinear_ssp_rk(distribution, basis, elliptic, grids, dg_flux, refs, energy_out=energy)
                self.check_finite(distribution, grids, dg_flux, elliptic)
                self.time += self.dt
                self.steps_counter += 1
                self.time_array = np.append(self.time_array, self.time)
                self.field_energy = np.append(self.field_energy, float(energy.item()))
            self.resident_end(distribution)
        else:
            if steps is None:
                steps = int(np.ceil((self.final_time - self.time) / self.dt - 1e-9))
            energies = torch.empty(steps + 1, dtype=torch.float64, device=dev)
            energies[0] = e0
            self.resident_begin(distribution, grids, dg_flux, elliptic)
            k = 0
            graph = None
            while k < steps:
                if graph_steps > 0 and k >= 1 and steps - k >= graph_steps:
                    if graph is None:
                        graph = StepGraph(self, distribution, basis, elliptic, grids, dg_flux, graph_steps, refs)
                    graph.replay()
                    energies[k + 1:k + 1 + graph_steps].copy_(graph.energies)
                    n_done = graph_steps
                else:
                    self.nonlinear_ssp_rk(distribution, basis, elliptic, grids, dg_flux, refs,
                                          energy_out=energies[k + 1:k + 2])
                    n_done = 1
                    if (k + 1) % self.CHECK_EVERY == 0:
                        self.check_finite(distribution, grids, dg_flux, elliptic)
                for _ in range(n_done):
                    self.time += self.dt
                    self.steps_counter += 1
                    self.time_array = np.append(self.time_array, self.time)
                k += n_done
            self.check_finite(distribution, grids, dg_flux, elliptic)
            self.resident_end(distribution)
            torch.cuda.synchronize(dev)
            self.field_energy = np.append(self.field_energy, energies.cpu().numpy())
        self.elapsed = _time.time() - t0

    def _main_loop_linear(self, distribution, basis, elliptic, grids, dg_flux, refs, steps, e0):
        """main/timestep.py:81-139 with the linear update of :189-229."""
        if self.fixed_dt is None:
            self._adapt(grids, elliptic, dg_flux, refs)
        self.field_energy = np.append(self.field_energy, float(e0.item()))
        done = 0
        while (self.time < self.final_time) if steps is None else (done < steps):
            done += 1
            self.ssp_rk_update(distribution, basis, grids, elliptic, dg_flux, refs)
            self.time += self.dt
            self.steps_counter += 1
            self.time_array = np.append(self.time_array, self.time)
            self.field_energy = np.append(self.field_energy, float(elliptic.electric_energy(grids.x).item()))

    def _main_loop_device_dt(self, distribution, basis, elliptic, grids, dg_flux, refs, steps, graph_steps, e0):
        """main/timestep.py:81-139 with the CFL rule on the device: the clock
        (dt, time) lives in the workspace; the host reads it once per graph
        replay (or per eager step) only to stop at ``final_time``.  A replay
        of K steps is used only when even K steps of the largest possible dt
        (max|acc E| = 0) stay below final_time, so the loop stops exactly
        where the reference's does."""
        import torch
        dev = distribution.arr.device
        self.field_energy = np.append(self.field_energy, float(e0.item()))
        self.resident_begin(distribution, grids, dg_flux, elliptic)
        g = self._res["geom"]
        ws = g.workspace(dev)
        g.set_cfl(self.courant, grids.x.dx, grids.u.dx, grids.v.dx)
        if elliptic.electric_field is None or elliptic.electric_field.device != dev:
            raise RuntimeError("main_loop: the set-up field must be solved first (main/main.py:131)")
        if elliptic.potential is None or elliptic.potential.device != dev:
            elliptic.potential = torch.zeros((grids.x.res, grids.x.order), dtype=torch.float64, device=dev)
        _lib.check(_lib.load().vl_clock_init(g.handle, _lib.ptr(elliptic.electric_field), float(self.time),
                                             _lib.ptr(ws), _lib.stream_ptr()), "vl_clock_init")
        # largest possible dt: the rule with max|acc E| = 0 (an upper bound of every step's dt)
        acc, b = dg_flux.acceleration_coefficient, elliptic.magnetic_field
        s0 = abs(grids.u.arr_max)
        bv = abs(acc * b * grids.v.arr_max)
        s2 = abs(acc * abs(b * grids.u.arr_max))
        dt_max = self.courant / ((s0 / grids.x.dx) + (bv / grids.u.dx) + (s2 / grids.v.dx))
        energies, clocks = [], []
        done = 0
        t_host = self.time
        graph = None
        K = int(graph_steps)
        while (done < steps) if steps is not None else (t_host < self.final_time):
            if steps is not None:
                use_graph = K > 0 and done >= 1 and steps - done >= K
            else:
                use_graph = K > 0 and done >= 1 and t_host + K * dt_max < self.final_time
            if use_graph:
                if graph is None:
                    graph = StepGraph(self, distribution, basis, elliptic, grids, dg_flux, K, refs, adaptive=True)
                graph.replay()
                energies.append(graph.energies.clone())
                clocks.append(graph.clocks.clone())
                done += K
            else:
                e = torch.empty(1, dtype=torch.float64, device=dev)
                c = torch.empty((1, 2), dtype=torch.float64, device=dev)
                self._fused_step(distribution, elliptic, grids, dg_flux, energy_out=e, clock_out=c[0], adaptive=True)
                energies.append(e)
                clocks.append(c)
                done += 1
            if steps is None:
                t_host = float(clocks[-1][-1, 0].item())
            if not use_graph and (steps is None or done % self.CHECK_EVERY == 0):
                self.check_finite(distribution, grids, dg_flux, elliptic)
        self.check_finite(distribution, grids, dg_flux, elliptic)
        self.resident_end(distribution)
        torch.cuda.synchronize(dev)
        if done:
            en = torch.cat(energies).cpu().numpy()
            ck = torch.cat(clocks).cpu().numpy()
            self.field_energy = np.append(self.field_energy, en)
            self.time_array = np.append(self.time_array, ck[:, 0])
            self.time = float(ck[-1, 0])
            self.dt = float(ck[-1, 1])
        self.steps_counter += done


class StepGraph:
    """``steps`` fixed-dt SSP-RK3 steps captured once in a HIP graph.

    The step sequence launches kernels only (no host synchronisation, no
    allocation, no memset), so it is captured as is on a side stream of
    default priority.  The first captured step reuses the moment partials the
    previous step left in the workspace, so the state must have been stepped
    once (eagerly) before capture; when a library call or a torch write has
    touched the workspace or the state since the last step, ``replay()``
    first recomputes those partials (``vl_moment_partials``).  ``replay()``
    advances the distribution by ``steps`` steps and leaves the energies of
    those steps in ``energies``.
    """

    def __init__(self, stepper, distribution, basis, elliptic, grids, dg_flux, steps, refs=None, adaptive=False):
        import torch
        if stepper.fixed_dt is None and not adaptive:
            raise ValueError("a captured step needs a fixed dt, or the device clock (adaptive=True)")
        if stepper._rho_token is None:
            raise RuntimeError("step once eagerly before capturing (the first captured step reuses "
                               "the moment partials of the previous step)")
        dev = distribution.arr.device
        self.steps = int(steps)
        self.energies = torch.zeros(self.steps, dtype=torch.float64, device=dev)
        self.clocks = torch.zeros((self.steps, 2), dtype=torch.float64, device=dev)   # (time, dt) per step
        side = torch.cuda.Stream(device=dev)
        side.wait_stream(torch.cuda.current_stream(dev))
        self.graph = _lib.HipGraph()
        # no garbage collection during capture: a collected geometry handle
        # would free device memory in the middle of it
        gc.collect()
        gc_was_enabled = gc.isenabled()
        gc.disable()
        try:
            with self.graph.capture(side):
                for k in range(self.steps):
                    if adaptive:
                        stepper._fused_step(distribution, elliptic, grids, dg_flux, energy_out=self.energies[k:k + 1],
                                            clock_out=self.clocks[k], adaptive=True)
                    else:
                        stepper.nonlinear_ssp_rk(distribution, basis, elliptic, grids, dg_flux, refs,
                                                 energy_out=self.energies[k:k + 1])
        finally:
            if gc_was_enabled:
                gc.enable()
        torch.cuda.current_stream(dev).wait_stream(side)
        self.stepper = stepper
        self._geom, self._arr, _, _, self._lay = stepper._step_state(distribution, grids, dg_flux, elliptic)
        self._ws = self._geom.workspace(dev)
        # the partials in the workspace are still those of the last eager step
        self._mark = Stepper._token(self._arr, self._ws)

    def replay(self, check=True):
        """Replay the captured steps; with ``check`` then read the workspace
        status (one synchronisation) and raise VlasovError if a step met a
        non-finite field."""
        if Stepper._token(self._arr, self._ws) != self._mark:
            _lib.check(_lib.load().vl_moment_partials(self._geom.handle, _lib.ptr(self._arr), _lib.ptr(self._ws),
                                                      self._lay, _lib.stream_ptr()), "vl_moment_partials")
        self.graph.launch()
        self._mark = Stepper._token(self._arr, self._ws)
        self.stepper._rho_token = self._mark
        if check:
            self.check()

    def check(self):
        self._geom.check_finite(self._arr.device, f"graph replay of {self.steps} SSP-RK3 steps")
