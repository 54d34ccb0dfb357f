"""FP64 CPU oracle for the kernel-convolution dose path.

TEST INFRASTRUCTURE ONLY.  Nothing in the shipped package imports this
module: only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline``
leg of ``bench.py`` use it, and only as the checker (or the timed CPU
restatement of the reference), never as the thing being measured.

Contents
--------
``oracle.conv``   restated convolution, time integration, epilogue, metrics
``oracle.dpk``    restated dose-point-kernel generators (Y-90, Lu-177, synthetic)
``oracle.brute``  ctypes wrapper of ``brute.c``: an independent plain-C
                  direct convolution used to pin the FFT restatement

Every function cites the reference ``file:line`` it restates.  How each is
pinned (closed forms, library routines, brute force) is listed in the
docstring of the module that holds it; no function here is "parity
unpinned".
"""
