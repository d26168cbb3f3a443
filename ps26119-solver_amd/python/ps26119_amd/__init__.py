"""Helpers around the MI355X r²HPDHG backend: LP generators (lpgen), the thin HIP layer (psh)."""
