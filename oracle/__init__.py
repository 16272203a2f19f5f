"""CPU oracle for the split-step wave-propagation path.

TEST INFRASTRUCTURE ONLY.  This package is a complex128 numpy/scipy
restatement of big-wave's in-RAM (`NumpyVector`) simulation path.  It exists
to check the HIP implementation; it is never part of the product path.  Only
`tests/`, `__graft_entry__.smoke()` and the `cpu_baseline` leg of `bench.py`
may import it, and only as the checker or the timed CPU baseline.

Modules
-------
bigwave  -- propagation, optical elements, sources, per-source driver and the
            sim-dir setup rules (RNG order, cutoff angles), following
            big-wave/{propagation,optical_element,source,wavesim,multisim,util}.py
xray     -- delta/beta of a material from Chantler form factors, following
            nist_lookup/nist_lookup/xraydb_plugin.py:134-180 and
            xraydb.py:371-403, reading a JSON fixture instead of SQLite.

Pinning (what fixes each function independently of this restatement) is
listed in each function's docstring; every function is pinned by at least one
known-answer test from the reference's own test suites or by the end-to-end
golden vector of big-wave/test.py:853-960.
"""
