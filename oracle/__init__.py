"""ORACLE -- test infrastructure, not product code.

CPU restatement of spinifel's M-TIP orientation-matching path, used only as
the checker (by ``tests/`` and ``__graft_entry__.smoke()``) and as the timed
CPU baseline (by ``bench.py``'s ``cpu_baseline`` leg).  The product path in
``spinifel_amd`` never imports anything from here.

Modules
-------
es_nufft   ES-kernel type-2 3D NUFFT (finufft's published method), FP64,
           interpolation in C/OpenMP (``oracle/esnufft.c``), plus the FP64
           direct NUDFT ground truth.
geometry   quaternion -> rotation -> inverse, rotated/normalised q-points.
matching   std scaling, sklearn-formula distances, argmin, and the
           ``slicing_and_match`` / ``gen_model_slices`` restatements.
"""
