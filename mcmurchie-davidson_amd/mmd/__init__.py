"""MI355X-native McMurchie-Davidson Hartree-Fock: the reference's
``mmd`` API (Molecule, SCF, integrals) over libmmd's HIP kernels."""
__version__ = "0.1.0"
