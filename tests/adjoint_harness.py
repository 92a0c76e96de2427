"""Check harness for the differentiable step: the impact parameters -> loss map on the device, its gradient
through the C-ABI reverse pass (isph_vjp_run), a torch.autograd.Function over it, and central differences.

Parameters (in this order): impact angle (degrees), impact speed (units of the mutual escape speed),
impactor / target mass ratio, and the granite mantle's Tillotson a, b, A, B (eos.rs:27-47).
"""
import numpy as np
import torch

from botheadstudios_amd import isph, materials, scene

PARAMS = ("angle_deg", "v_over_vesc", "mass_ratio", "a", "b", "A", "B")
_EOS_COLS = (1, 2, 3, 4)  # a, b, A, B in the f64 Tillotson row


def nominal():
    g = materials.eos_f64()[materials.MAT_GRANITE]
    return np.array([45.0, 1.2, 0.11, g[1], g[2], g[3], g[4]])


class ImpactProblem:
    """Loss after `steps` KDK steps of the seeded two-body scene, as a function of the parameters. The dt
    sequence is taken from the first gradient evaluation (Courant at CFL 0.3) and then held fixed, as the
    reverse pass holds it fixed."""

    def __init__(self, n_target, n_impactor, seed, steps, checkpoint_every=10, gravity=isph.GRAVITY_DIRECT,
                 r_cut=2.9, k_r=20.0, k_j=10.0, cfl=0.3):
        self.family = scene.ImpactFamily(n_target, n_impactor, seed)
        self.n = n_target + n_impactor
        self.steps, self.K, self.gravity, self.cfl = steps, checkpoint_every, gravity, cfl
        self.loss = isph.Loss.make(scene.L_U, r_cut=r_cut, k_r=k_r, k_j=k_j)
        self.dts = None
        self.last = None

    def sim(self, p):
        rec, pos, vel = self.family.state(p[0], p[1], p[2])
        eos64 = materials.eos_f64().copy()
        eos64[materials.MAT_GRANITE, list(_EOS_COLS)] = p[3:7]
        s = isph.GpuSph(self.n)
        s.upload(rec, materials.eos_table(), self.family.softening * scene.L_U)
        s.set_eos_f64(eos64)
        s.set_precision(isph.PRECISION_MIXED)
        s.set_state_f64(pos, vel)
        s.set_gravity(self.gravity, 0.5, 1)
        return s, rec

    def gradient(self, p):
        """(loss, d loss / d params [7]) through the reverse pass; also keeps the raw outputs in self.last."""
        s, _ = self.sim(p)
        with s:
            out = s.vjp_run(self.loss, self.steps, self.K, cfl=self.cfl, dts=self.dts)
        if self.dts is None:
            self.dts = out["dts"].copy()
        g = np.zeros(7)
        for k, (dpos, dvel, dm) in enumerate(self.family.tangents(p[0], p[1], p[2])):
            g[k] = (out["dpos"] * dpos).sum() + (out["dvel"] * dvel).sum() + (out["dmass"] * dm).sum()
        g[3:7] = out["deos"][materials.MAT_GRANITE]
        self.last = out
        return out["loss"], g

    def value(self, p, dpos=None, dvel=None):
        """Loss with the fixed dt sequence, optionally from a perturbed f64 start state."""
        assert self.dts is not None, "run gradient() first: it fixes the dt sequence"
        s, rec = self.sim(p)
        with s:
            if dpos is not None or dvel is not None:
                _, pos, vel = self.family.state(p[0], p[1], p[2])
                s.set_state_f64(pos + (0 if dpos is None else dpos), vel + (0 if dvel is None else dvel))
            for dt in self.dts:
                s.set_dt(float(dt), 1.0)
                s.encode_kdk(1)
            return s.loss_value(self.loss)

    def realized_step(self, p, k, h):
        """The parameter step actually applied: masses go through f32 records, so a mass-ratio step is
        measured on the rounded impactor mass."""
        if k != 2:
            return h
        m_plus = np.float32(self.family.state(p[0], p[1], p[2] + h)[0]["mass"][-1])
        m_minus = np.float32(self.family.state(p[0], p[1], p[2] - h)[0]["mass"][-1])
        dm_dq = self.family.m_t / self.family.n_impactor * scene.M_U
        return (float(m_plus) - float(m_minus)) / (2.0 * dm_dq)

    def central_differences(self, p, rel=1e-4, rel_eos=1e-3):
        """Central differences of the device forward run. The Tillotson parameters take a larger step: the
        forward evaluates its pair terms at f32-stored coordinates, so the 1e-4 responses of these weak
        parameters sit at the f32 resolution of the positions they move."""
        g = np.zeros(7)
        for k in range(7):
            h = (rel if k < 3 else rel_eos) * abs(p[k])
            pp, pm = p.copy(), p.copy()
            pp[k] += h
            pm[k] -= h
            g[k] = (self.value(pp) - self.value(pm)) / (2.0 * self.realized_step(p, k, h))
        return g


class ImpactLoss(torch.autograd.Function):
    """torch view of the device loss: forward runs the reverse pass as well and keeps the gradient."""

    @staticmethod
    def forward(ctx, theta, problem):
        loss, grad = problem.gradient(theta.detach().cpu().numpy().astype(np.float64))
        ctx.save_for_backward(torch.from_numpy(grad))
        return theta.new_tensor(loss)

    @staticmethod
    def backward(ctx, g):
        (grad,) = ctx.saved_tensors
        return g * grad.to(g.dtype), None
