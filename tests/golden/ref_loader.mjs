// ESM loader hook for Node 12: the reference's sources are ES modules with a
// .js extension and no package.json, so files under the directory named by
// GEON_REF_SRC are loaded as modules.  Used only by make_reference_goldens.mjs,
// in the build container, to write fixtures.
const base = process.env.GEON_REF_SRC ? new URL('file://' + process.env.GEON_REF_SRC.replace(/\/?$/, '/')).href : null;

export async function getFormat(url, context, defaultGetFormat) {
    if (base && url.startsWith(base) && url.endsWith('.js')) return { format: 'module' };
    return defaultGetFormat(url, context, defaultGetFormat);
}
