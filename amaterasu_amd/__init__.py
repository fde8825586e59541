"""amaterasu_amd: the plane-wave H.psi / Davidson / density hot path of amaterasu on MI355X.

Drop-in entry points with the reference call signatures:

* :class:`Hamiltonian`      -- ``amaterasu.hamiltonian.operator.Hamiltonian``
* :func:`solve`, :func:`davidson`, :class:`DavidsonResult` -- ``amaterasu.scf.eigen``
* :func:`compute_density`   -- ``amaterasu.scf.density.compute_density``
* :class:`FFTGrid`          -- ``amaterasu.basis.fft.FFTGrid`` (transforms on the GPU)

The compute runs in the in-tree HIP library ``libamaterasu_amd.so`` (gfx950); importing
works without a GPU, computing does not.
"""

from ._lib import AmtError

__version__ = "0.1.0"


def __getattr__(name):
    # Lazy imports keep ``import amaterasu_amd`` cheap and torch-free until needed.
    if name in ("Hamiltonian", "device_plan"):
        from . import hamiltonian

        return getattr(hamiltonian, name)
    if name in ("davidson", "solve", "DavidsonResult", "random_bands", "orthonormalize", "rayleigh_ritz",
                "orthogonalize_against", "teter_payne_allan"):
        from . import eigen

        return getattr(eigen, name)
    if name in ("compute_density", "electron_count"):
        from . import density

        return getattr(density, name)
    if name in ("FFTGrid", "GSphere"):
        from . import basis

        return getattr(basis, name)
    raise AttributeError(name)


__all__ = ["AmtError", "Hamiltonian", "DavidsonResult", "FFTGrid", "GSphere", "compute_density", "davidson",
           "solve"]
