"""MI355X-native constant-density acoustic time solver with PySIT's API."""
