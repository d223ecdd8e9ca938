"""``python -m toptimization_amd`` (reference ``__main__.py``)."""

import sys

from .cli import main

sys.exit(main())
