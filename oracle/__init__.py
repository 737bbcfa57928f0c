"""CPU oracle for the VBR SpMV/SpMM hot path — TEST INFRASTRUCTURE ONLY.

Nothing in the product (``sable_amd``) may import, call, link or execute
anything under ``oracle/``.  Only ``tests/``, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` use it, and there only as the checker
(or as the timed CPU baseline), never as the thing measured on the GPU.

Contents
--------
``vbr_oracle``  pure-Python loop restatements of the reference's inspector
                (VBR conversion, block routing, CSR assembly, dense expansion,
                partition cut points) and an FP64 SpMV/SpMM over a VBR.
``spmv_ref.c``  a C restatement of the generated-C executor semantics
                (CSR part first, then dense blocks in ordinal order with the
                three kernel shapes; SpMM with the RHS width as a parameter),
                built by ``oracle/Makefile`` into ``oracle/_build/``.
``fileio_oracle``  the reference's text parse (Python float()/int() per
                token), the baseline and checker of the native text reader.
``ref_build``   recipe that compiles the reference's own ``src/csr-spmv.cpp``
                from ``/root/reference`` into ``oracle/_ref/`` when present.

Pin status of each function is listed in the module docstrings.
"""
