"""MI355X (gfx950) implementation of libprofit's image-generation path.

The computation lives in the HIP library libprofit_amd/_lib/libprofit_hip.so
behind the C-ABI of include/profit_hip.h; this package only mirrors
libprofit's Model / Profile / Convolver interface on top of it.
"""
from . import abi  # noqa: F401

__version__ = "1.9.3"


def __getattr__(name):
    # Model, Profile, Context, create_convolver are loaded lazily so that
    # `import libprofit_amd` works (and fails loudly on use) without the
    # library.
    if name in ("Model", "Profile", "Context", "Convolver", "create_convolver", "default_context", "evaluate_spec",
                "Batch", "init", "finish", "clear_cache", "evaluate_batch", "chi2_grad_once"):
        from . import model
        return getattr(model, name)
    raise AttributeError(name)
