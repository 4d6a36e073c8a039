"""Oracle package: CPU restatement of the reference step, used only as the
checker in tests, smoke() and as the timed CPU baseline in bench.py."""
