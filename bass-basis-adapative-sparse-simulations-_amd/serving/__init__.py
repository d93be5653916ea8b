"""Batched and served execution of many BASS circuits on one GPU."""
