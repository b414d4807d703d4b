"""The reference-side C callers shown in INTEGRATION.md (sections 2a-2e and
4) compile against include/kssdhip.h and link against libkssdhip.so with no
undefined symbol.  The reference's own declarations the snippets use (its
types, globals and the variables of the functions they sit in) are given
here as stand-ins with the reference's types (command_ani.h:33-45,
:114-124; sketch_rearrange.h:19); each snippet becomes the body of a
function of its own."""
import re
import shutil
import subprocess

import pytest

import kssdtest as kt

DOC = kt.ROOT / "INTEGRATION.md"

PRELUDE = r"""
#include <err.h>
#include <stdbool.h>
#include <stdint.h>
#include <stdlib.h>
#include "kssdhip.h"

typedef struct __attribute__((packed)) { uint64_t ctxgid; uint32_t obj; } ctxgidobj_t;
typedef struct { uint32_t num_ctx, num_mut2_ctx; } ctx_mut2_t;
typedef struct { uint32_t diff_obj, diff_obj_section; } obj_section_t;
typedef struct { int ntop; int ctxcut; float afcut; bool afcut_set; float anicut; bool ignoreconflict; } ani_opt_t;
typedef struct { uint32_t index_max_ctx_freq, index_sample_step; } pairwise_opt_t;
typedef struct { uint64_t *comb_sketch; uint64_t *sketch_index; int infile_num; bool conflict; } sketch_result_t;
struct { unsigned ctx, obj; } Bitslen;
int ani_model_drfold;
ctxgidobj_t *sortedcomb_ctxgid64obj32;
size_t ref_sketch_size, sketch_size, rows_cap;
int ref_infile_num, this_block_size, offset_gid, this_block, offset, nb, n, q0;
ani_opt_t *ani_opt;
pairwise_opt_t *opt;
sketch_result_t *ref_result;
ctx_mut2_t *ctx, *unique_ctx;
obj_section_t *obj, *unique_obj;
uint64_t *tmp_ctxobj, *this_sketch_index, *keys, *skidx, *comb_sketch, *sketch_index, *arr_qry, *arr_ref;
uint32_t len_qry, len_ref, flags, *qry_ctx_num, *ref_ctx_num;
kssd_ctx_mut2_t *ctx_row;
kssd_obj_section_t *obj_row;
kssdhip_ani_row_t *rows;
kssdhip_index *gpu_ix;
kssdhip_serve_stats_t stats;
"""


def snippets():
    text = DOC.read_text()
    out = []
    for sec in re.split(r"^(?=#{2,3} )", text, flags=re.M):
        head = sec.splitlines()[0] if sec else ""
        if not re.match(r"#{2,3} (2[a-e]\.|4\. Serving)", head):
            continue
        for i, body in enumerate(re.findall(r"```c\n(.*?)```", sec, re.S)):
            out.append((f"{head.split()[1]}{i}", body))
    return out


def test_snippets_found():
    names = [n for n, _ in snippets()]
    assert len(names) >= 7, names
    assert any(n.startswith("2a") for n in names) and any(n.startswith("4.") for n in names)


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_snippets_compile_and_link(tmp_path):
    src = [PRELUDE]
    for k, (name, body) in enumerate(snippets()):
        lines = [ln for ln in body.splitlines() if not ln.startswith("#include")]
        src.append(f"/* {name} */\nvoid snippet_{k}(void)\n{{\n" + "\n".join(lines) + "\n}\n")
    c = tmp_path / "callers.c"
    c.write_text("\n".join(src))
    base = ["gcc", "-std=gnu11", "-Wall", "-Werror", "-Wno-unused-variable", "-Wno-unused-but-set-variable",
            "-I", str(kt.ROOT / "include")]
    p = subprocess.run([*base, "-c", "-fPIC", "-o", str(tmp_path / "callers.o"), str(c)], capture_output=True,
                       text=True)
    assert p.returncode == 0, p.stderr + "\n" + c.read_text()
    lib = kt.PKG / "libkssdhip.so"
    if not lib.exists():
        pytest.skip("libkssdhip.so not built")
    p = subprocess.run(["gcc", "-shared", "-o", str(tmp_path / "callers.so"), str(tmp_path / "callers.o"),
                        f"-L{kt.PKG}", "-lkssdhip", "-Wl,--no-undefined", "-Wl,--allow-shlib-undefined"],
                       capture_output=True, text=True)
    assert p.returncode == 0, p.stderr
