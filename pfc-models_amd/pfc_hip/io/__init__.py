"""Job keys and result formats shared with the reference's tooling."""
from .runid import config_hash, run_id  # noqa: F401
