"""ORACLE — CPU restatement of madflow's hot path, used only as the checker by tests/,
the smoke check and bench.py's CPU baseline.  Never imported by the product package."""
