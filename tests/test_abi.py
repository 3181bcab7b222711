"""CPU checks of the C ABI: the library loads, exports every function that
include/aw.h declares, and the ctypes binding covers all of them.  No
compute call is made (there may be no GPU here)."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


def declared_functions():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        for m in re.finditer(r"^\s*(?:aw_status|void|const char\s*\*)\s+(aw_\w+)\s*\(", text,
                             flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_functions()
    assert "aw_sim_step" in names and "aw_push_deposit_es" in names
    assert len(names) >= 30


def test_library_exports_every_declared_symbol():
    from arcwarden_amd import _lib
    if not _lib.LIB_PATH.exists():
        pytest.fail("libarcwarden_hip.so is not built (run __graft_entry__.build())")
    L = ctypes.CDLL(str(_lib.LIB_PATH))
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, f"not exported: {missing}"


def test_binding_covers_header():
    from arcwarden_amd import _lib
    assert set(declared_functions()) == set(_lib.exported_symbols())


def test_struct_layouts_match_header_sizes():
    """The ctypes mirrors must have the C sizes (compiled probe of aw.h)."""
    import subprocess
    import tempfile
    from arcwarden_amd import _lib
    src = r"""
#include <stdio.h>
#include "aw.h"
int main(void){printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(aw_grid),
 sizeof(aw_run_params), sizeof(aw_species), sizeof(aw_particles), sizeof(aw_fields),
 sizeof(aw_sources), sizeof(aw_diag_sample), sizeof(aw_sim_config), sizeof(aw_deck),
 sizeof(aw_run_desc), sizeof(aw_ensemble_stats));
 return 0;}
"""
    with tempfile.TemporaryDirectory() as d:
        c = Path(d) / "probe.c"
        c.write_text(src)
        exe = Path(d) / "probe"
        subprocess.run(["gcc", "-I", str(ROOT / "include"), str(c), "-o", str(exe)], check=True)
        sizes = [int(v) for v in subprocess.run([str(exe)], capture_output=True,
                                                text=True).stdout.split()]
    py = [ctypes.sizeof(t) for t in (_lib.AwGrid, _lib.AwRunParams, _lib.AwSpecies,
                                     _lib.AwParticles, _lib.AwFields, _lib.AwSources,
                                     _lib.AwDiagSample, _lib.AwSimConfig, _lib.AwDeck,
                                     _lib.AwRunDesc, _lib.AwEnsembleStats)]
    assert py == sizes
