"""CPU oracle of the shoebox-integration path -- TEST INFRASTRUCTURE ONLY.

This package restates, in float64 numpy, the algorithm of XPG's profile-fitting integration kernels
(xpg/integrate_kernels.py, xpg/shoebox.py and the box planning / profile learning parts of xpg/integrate.py).
It is the checker the HIP kernels are compared against.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it; the product package ``xpg_amd`` never does.

Pinning (how each function is checked before it is trusted) is listed in the header of each module.
"""
