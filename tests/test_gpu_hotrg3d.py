"""HOTRG_3D on the device (src/schemes/hotrg3d.jl:48-141) through the C ABI:
per-step norms against the oracle (rtol 1e-10) on Z2 and Trivial 3D Ising
tensors, the reference's own free-energy test, and the 6-leg tensor format
crossing the ABI unchanged."""
import numpy as np
import pytest

from oracle import models as om
from oracle import schemes as osch

pytestmark = pytest.mark.gpu


# Cases whose truncations all have a gap at the cut (the oracle's log): at
# beta 0.25, chi 8 (and beta_c, chi 8) the cut splits exactly degenerate
# eigenvalues, so the kept subspace is not a function of the tensor there.
@pytest.mark.parametrize("sym,beta,chi,n", [("Z2", None, 4, 3), ("Z2", None, 6, 3),
                                            ("Trivial", 0.2, 4, 2), ("Z2", 0.25, 6, 3),
                                            ("Z2", 0.3, 6, 3)])
def test_hotrg3d_vs_oracle(tk, sym, beta, chi, n):
    b = om.ising_betac_3d if beta is None else beta
    T = om.classical_ising_3d_z2(b) if sym == "Z2" else om.classical_ising_3d_trivial(b)
    ref = np.array(osch.run(osch.HOTRG_3D(T), chi, n))
    sc = tk.HOTRG_3D(tk.classical_ising_3D(beta, sym))
    got = np.array(tk.run(sc, tk.truncrank(chi), tk.maxiter(n), verbosity=0))
    sc.close()
    assert np.max(np.abs(got / ref - 1)) < 1e-10


def test_hotrg3d_reference_gate(tk):
    """test/schemes.jl:376-383 on the device: chi 8, 25 steps, f within 1e-3
    of -3.507."""
    sc = tk.HOTRG_3D(tk.classical_ising_3D())
    d = tk.run(sc, tk.truncrank(8), tk.maxiter(25), verbosity=0)
    sc.close()
    f = tk.free_energy(d, tk.ising_betac_3D, scalefactor=8.0)
    assert abs(f / tk.f_benchmark3D - 1) < 1e-3


def test_hotrg3d_tensor_crosses_the_abi(tk):
    T = tk.classical_ising_3D(0.3, "Z2")
    sc = tk.HOTRG_3D(T)
    back = sc.T
    assert np.array_equal(back.data, T.data) and np.array_equal(back.leg_dims, T.leg_dims)
    tk.step(sc, tk.truncrank(4))
    t2 = sc.T
    sc.T = t2
    assert np.array_equal(sc.T.data, t2.data)
    sc.close()
