"""Backend registration (mirrors `src/gpufsm/backends/__init__.py:18-39`).

Importing this package registers every backend subpackage. A subpackage that fails to
import is recorded in :data:`IMPORT_ERRORS` instead of vanishing silently.
"""

from __future__ import annotations

import importlib

IMPORT_ERRORS: dict[str, str] = {}

for _backend in ("gpufsm.backends.hip",):
    try:
        importlib.import_module(_backend)
    except Exception as _exc:  # pragma: no cover - depends on the build
        IMPORT_ERRORS[_backend.rsplit(".", 1)[-1]] = f"{type(_exc).__name__}: {_exc}"
