"""MI355X-native divide-and-conquer tridiagonal eigensolver (HSSPACK mdstedc path)."""
