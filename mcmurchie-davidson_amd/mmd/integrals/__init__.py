"""Integral entry points with the reference's module layout:
``mmd.integrals.twoe`` (Basis, ERI, doERIs), ``mmd.integrals.fock`` (formPT)
and ``mmd.integrals.onee`` (S, T, V, Mu, RxDel, _boys).  All evaluation runs
in libmmd's HIP kernels."""
