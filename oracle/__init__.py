"""CPU oracle of the cortical cycle -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
package; the engine (bismit_amd) never does.
"""
