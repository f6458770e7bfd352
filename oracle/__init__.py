"""TEST INFRASTRUCTURE: CPU oracle of the facet gridder / degridder.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this package,
and only as the checker or as the timed CPU baseline -- never on the product path.

    cpuref   ctypes wrapper of cpu_ref.cc, the C++ restatement of the reference gridder
             (DDFacet/Gridder/gridder.h, degridder.h and helpers)
    refmod   loader of the reference module built from its own sources into oracle/_ref/
             (make -f oracle/Makefile.ref), used to pin cpuref and write golden vectors
    dft      FP64 direct Fourier transforms of point-source skies (physics oracle)
"""
