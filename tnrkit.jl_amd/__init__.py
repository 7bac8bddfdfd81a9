"""MI355X engine for TNRKit.jl's TRG / BTRG / HOTRG coarse-graining loop.

The arithmetic runs in libtnr_hip.so (HIP, gfx950); this package is the host
side mirroring TNRKit's interface:

    T = classical_ising(symmetry="Z2")          # beta_c
    scheme = BTRG(T)
    data = run(scheme, truncrank(16), maxiter(25))
    f = free_energy(data, ising_betac)

The directory name is not a valid Python identifier; load it with
``__graft_entry__.package()`` (it registers the module as ``tnrkit_jl_amd``).
"""
from ._lib import TNRError, lib  # noqa: F401
from .models import (classical_clock, classical_ising, classical_ising_3D,  # noqa: F401
                     classical_ising_impurity, f_benchmark3D, ising_betac_3D,
                     classical_potts, f_onsager,
                     ising_betac, ising_cft_exact, potts_betac)
from .schemes import (ATRG, BTRG, FP64, HOTRG, HOTRG_3D, SPLIT32, TRG, Finalizer, ImpurityHOTRG,  # noqa: F401
                      ImpurityTRG, MultipleCrit,
                      convcrit, default_Finalizer, finalize, free_energy, maxiter, run, solver_config, step,
                      trivial_convcrit, truncrank)
from .observables import (GSDegeneracy_Finalizer, central_charge, cft_data,  # noqa: F401
                          finalize_central_charge, finalize_groundstatedegeneracy,
                          finalize_gu_wen_ratio, finalize_two_by_two, ground_state_degeneracy,
                          gu_wen_ratio, guwenratio_Finalizer, transfer_spectrum,
                          two_by_two_Finalizer)
from .serving import Request, Server, run_batch  # noqa: F401
from .tensors import NTensor, SymTensor, trivial  # noqa: F401
