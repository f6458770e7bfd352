"""MI355X-native facet gridding / degridding beneath DDFacet's ClassDDEGridMachine.put()/get()."""
__version__ = "0.1.0"
