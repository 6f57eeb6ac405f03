"""MI355X (gfx950) all-pairs entropy / mutual-information engine for PARENT trajectories.

Python mirror of the engine's C ABI (``include/parent_hip.h``), loaded with
ctypes from the in-tree ``parent_amd/_build/libparent_hip.so``.  There is no
CPU fallback: if the library is missing, importing the API raises.

The reference (markusfleck/PARENT) exposes this path as the programs
``PARENT.x`` (PARENT.cpp) and ``get_PAR_MIST.x`` (get_PAR_MIST.cpp); the
functions here follow them:

* :func:`run_parent` -- ``PARENT.x in.bat out.par b1 a1 d1 b2 a2 d2``
* :func:`run_mist`   -- ``get_PAR_MIST.x in.par out.txt``
* :func:`mie_write`  -- ``get_PAR_MIE.x in.par > out_MIE.txt``
* :func:`get_mutual` -- ``get_mutual`` (util.cpp:47-81)
"""
from .api import (  # noqa: F401
    METHOD_AUTO,
    METHOD_ATOMIC,
    METHOD_MFMA,
    METHOD_PLAIN,
    METHOD_SLABS,
    Bat,
    Batch,
    bat_build,
    Engine,
    ParentError,
    dof_counts,
    get_mutual,
    lib,
    library_path,
    mie_write,
    par_pos,
    read_par,
    run_mist,
    run_parent,
    synth_values,
    synth_write_bat,
    write_mist_text,
)
