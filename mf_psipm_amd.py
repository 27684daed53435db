"""Import shim: makes the ``mf-psipm_amd/`` directory importable as ``mf_psipm_amd``.

The package directory name carries a hyphen, which Python's import statement
cannot spell.  Importing this module loads ``mf-psipm_amd/__init__.py`` as the
package ``mf_psipm_amd`` and replaces this shim in ``sys.modules``.
"""
import importlib.util
import pathlib
import sys

_DIR = pathlib.Path(__file__).resolve().parent / "mf-psipm_amd"
_spec = importlib.util.spec_from_file_location(
    __name__, _DIR / "__init__.py", submodule_search_locations=[str(_DIR)]
)
_module = importlib.util.module_from_spec(_spec)
sys.modules[__name__] = _module
_spec.loader.exec_module(_module)
