"""Process exit with live handles: problems, plans and contexts still
referenced (module globals, a reference cycle, the frames of a caught
exception's traceback) when the interpreter shuts down.  The binding
destroys them in an atexit handler while the runtime is intact; the process
must end with status 0 (a crash at this point was seen after a failed test
under pytest -x)."""
import os
import subprocess
import sys
import textwrap

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SCRIPT = textwrap.dedent("""
    import sys
    sys.path.insert(0, %(root)r)
    %(setup)s
    import sparselizard_amd as sa
    ctx = %(ctx)s
    kept = sa.Problem(ctx, "poisson2d", "n=6")
    kept.plan()
    %(assemble)s
    class Node:
        pass
    a, b = Node(), Node()
    a.other, b.other = b, a
    a.prob = sa.Problem(ctx, "elasticity3d", "n=2")
    a.prob.plan()
    def fail(p):
        raise RuntimeError("held by the traceback")
    try:
        fail(sa.Problem(ctx, "poisson2d", "n=3"))
    except RuntimeError:
        err = sys.exc_info()
    del a, b
    print("exit with live handles", flush=True)
""")


def run_exit_script(ctx_expr, setup, assemble):
    code = SCRIPT % {"root": ROOT, "ctx": ctx_expr, "setup": setup, "assemble": assemble}
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr[-2000:])
    return r


def test_exit_with_live_host_plans():
    # pattern-only plans (no device context)
    r = run_exit_script("None", "", "")
    assert r.returncode == 0 and "exit with live handles" in r.stdout


@pytest.mark.gpu
def test_exit_with_live_device_plans():
    # device plans with assembled matrices, a torch allocation alive as well
    r = run_exit_script("sa.Context(0)", "import torch\ntorch.cuda.init()\nbuf = torch.ones(1024, device='cuda')",
                        'kept.generate("f16split")')
    assert r.returncode == 0 and "exit with live handles" in r.stdout
