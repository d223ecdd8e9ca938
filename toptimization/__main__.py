"""``python -m toptimization`` (reference ``__main__.py:1-6``)."""

import sys

from toptimization_amd.cli import main

sys.exit(main())
