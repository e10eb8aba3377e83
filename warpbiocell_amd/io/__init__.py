"""Run output in the reference's file formats."""
