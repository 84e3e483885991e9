"""CPU oracle for the SPECT system-matrix path (test infrastructure only).

This package restates PyTomography's rotate+sum SPECT projector, its
attenuation and PSF transforms, the Poisson likelihood, OSEM/BSREM and the
relative-difference prior as plain torch-CPU (and numpy brute-force) code.

It is the *checker*: only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it.  The product package
``pytomography_amd`` never imports anything from here.
"""
